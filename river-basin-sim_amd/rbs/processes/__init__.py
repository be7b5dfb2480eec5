"""Slow and point processes around the fast step -- reference solver/processes."""
