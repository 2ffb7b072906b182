"""The reference's ``utils`` package surface, backed by libdiver."""
