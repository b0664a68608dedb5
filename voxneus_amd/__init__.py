"""MI355X-native VoxNeuS voxel-SDF rendering and regularisation path."""
