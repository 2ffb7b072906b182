"""The reference's ``model`` package surface."""
