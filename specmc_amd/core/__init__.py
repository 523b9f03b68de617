"""Driver-level helpers of the reference's ``specmc.core`` that the public
surface uses: the CFL policy and the benchmark initial conditions."""
