"""Scenario config, status file and result store -- reference solver/io contracts."""
