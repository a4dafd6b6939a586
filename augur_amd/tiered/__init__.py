"""Mirror of the reference's model/tiered interfaces on the PM path."""
