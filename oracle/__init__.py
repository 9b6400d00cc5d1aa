"""CPU oracle package for the batched-FIR path (test infrastructure only;
the product never imports it)."""
