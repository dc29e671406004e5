"""compact3d_amd: MI355X-native per-timestep hot path of the compact3d channel solver."""
