"""MI355X-native time-stepping core for FD3D_TSN dynamic rupture simulations."""
