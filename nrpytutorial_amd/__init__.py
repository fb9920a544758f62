"""MI355X-native finite-difference RHS + Method-of-Lines engine with NRPy+'s codegen API."""
