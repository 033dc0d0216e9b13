"""deeprecon API mirror (MI355X build)."""
