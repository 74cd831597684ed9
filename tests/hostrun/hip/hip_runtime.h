// Empty: tests/hostrun/shim.h stands in for the HIP runtime in the host build.
