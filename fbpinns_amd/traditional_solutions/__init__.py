"""Reference solutions used by Problem.exact_solution for L1 errors."""
