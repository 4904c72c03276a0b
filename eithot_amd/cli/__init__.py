"""Command-line drivers with the reference's argv and output (test_inplace,
Nd_estimation) and the benchmark workflow over them."""
