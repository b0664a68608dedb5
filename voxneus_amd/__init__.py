"""MI355X-native VoxNeuS voxel-SDF rendering and regularisation path.

The public names mirror the reference's ``voxneus`` package (models, rendering,
geometry, regularization, training.scheduler, training.optimizer), so the hot path can be switched by import.  The
compute runs in hand-written gfx950 kernels (``libvoxneus_hip.so``, C ABI in
``include/voxneus_hip.h``); importing a name that needs the library raises if it has not
been built.
"""

__version__ = "0.1.0"

_LAZY = {
    "VoxNeuS": "voxneus", "DenseSDFGrid": "sdf_grid", "get_voxel_corners_and_weights": "sdf_grid",
    "interpolate_gradients": "sdf_grid", "world_to_grid_index": "sdf_grid",
    "RadianceHashGrid": "hash_grid", "PyTorchHashGrid": "hash_grid", "HashGridEncoder": "hash_grid",
    "RadianceMLP": "radiance", "SphericalHarmonicsEmbedder": "radiance", "SingleVarianceNetwork": "radiance",
    "PositionalEmbedder": "radiance",
    "RaySampler": "rendering", "OccupancyGrid": "rendering", "compute_neus_alpha": "rendering",
    "composite_weights": "rendering", "volume_render_from_samples": "rendering",
    "compute_huber_color_loss": "rendering", "near_far_from_sphere": "training", "get_rays": "training",
    "extract_active_vertices": "regularization", "compute_eikonal_loss_and_gradient": "regularization",
    "compute_curvature_loss_and_gradient": "regularization",
    "amplify_s_gradient": "training", "FusedTrainStep": "training", "StepConfig": "training",
    "FusedOptimizer": "training", "SuperResolutionSchedule": "training", "compute_analytical_gradient": "sdf_grid",
    "SparseGridAdam": "optim", "NetAdam": "optim", "create_optimizer": "optim",
    "upsample_sdf_grid": "sdf_grid", "get_resolution_for_step": "sdf_grid", "logistic_density": "rendering",
    "extract_mesh_marching_cubes": "mesh", "marching_cubes": "mesh", "apply_gaussian_filter_to_grid": "mesh",
    "VoxNeuSScheduler": "scheduler", "compute_neus2_lambda": "scheduler", "FusedTrainer": "trainer",
    "save_checkpoint": "checkpoint", "load_checkpoint": "checkpoint", "load_resident_scene": "checkpoint",
}


def __getattr__(name):
    mod = _LAZY.get(name)
    if mod is None:
        raise AttributeError(name)
    import importlib

    return getattr(importlib.import_module(f"{__name__}.{mod}"), name)
