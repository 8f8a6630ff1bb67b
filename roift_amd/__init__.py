"""roift_amd: MI355X-native OIFT (f_max) + fuzzy relaxation."""
