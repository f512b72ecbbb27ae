"""Collision steps: drop-ins of the reference's collision package."""
