"""CPU oracle for the Grassmann-tensor hot path -- test infrastructure only (see gtn_oracle.py)."""
