"""CPU oracle for the FEM hot path (test infrastructure only; see rb_oracle.py)."""
