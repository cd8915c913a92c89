"""MI355X-native Yee-FDTD engine for the lens_sensor waveguide solvers."""
