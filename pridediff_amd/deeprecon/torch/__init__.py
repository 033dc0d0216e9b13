"""deeprecon.torch API mirror."""
