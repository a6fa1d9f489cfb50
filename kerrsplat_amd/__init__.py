"""kerrsplat_amd: MI355X (gfx950) backend of KerrSplat's fused geodesic march + polarized transfer.

The host side mirrors the reference's Julia API (src/Geodesics/cache.jl, src/Splats/*.jl) over the
C-ABI of libkerrsplat_hip.so (include/kerrsplat_hip.h):

    cam   = Camera((-12, 12), (-12, 12), 64)
    cache = GeodesicCache(cam, N=512, store_samples=False)
    regenerate(cache, a, theta_o, marcher=Fused(64))
    img   = thin_image(cache, params, t_obs)                       # screen-shaped Stokes I
    st    = polarized_image(cache, params, t_obs, nu, L)          # (n_alpha, n_beta, 4)
    cube  = polarized_cube(cache, params, times, nus, L)          # (n_alpha, n_beta, nt, nnu, 4)
    dp, cube = polarized_gradient(cache, params, times, nus, L, dstokes)   # dual sweep (stored samples)
    chi = fit.chi2_gradient(dp, params, StokesMovie(...), cache, L)       # χ² and its gradient, seeds on the device

Every call goes to the GPU library; there is no CPU fallback. Loading fails loudly when the library
is missing or no GPU is present.
"""
from .api import (Camera, Direct, Recurrence, Fused, GeodesicCache, Model, CompositeModel, regenerate, thin_image, polarized_image,
                  polarized_cube, polarized_gradient, march, precision, stage_ms, render_device, visibilities, unsort, to_screen, samples, residuals, consts, KernelError,
                  NSPLATPARAMS, NPOLARIZEDPARAMS, NPOWERLAWPARAMS, NKAPPAPARAMS, gravitational_radius, grid_axis,
                  library_path, gradient_forward, gradient_reverse, vis_chi2_seed, image_chi2_seed, vis_seed, closure_chi2_seed, tiles, Binning,
                  bin, binned_grid, binned_polar, concatenate, spacetime_tangent, spacetime_tangent_differences, RayLists, ray_lists, framed, flattened, frame_lists,
                  concatenate_screens, screen_tiles, tuning, autotune, queue_costs, FitGraph, march_precision)
from . import fit
from .fit import (ScatteringKernel, taper, spacetime_valgrad, spacetime_jacobian, spacetime_movie_loss, StokesMovie, FrameScans, chi2, chi2_gradient, sweep_passes, frame_chi2_seed, image_loss_gradient, closure_phases,
                  log_closure_amplitudes, ClosureData, closure_chi2, frame_chi2_closures, gain_factors, apply_gains)

__all__ = ["Camera", "Direct", "Recurrence", "Fused", "GeodesicCache", "Model", "CompositeModel", "regenerate", "thin_image",
           "polarized_image", "polarized_cube", "polarized_gradient", "march", "precision", "stage_ms", "render_device", "visibilities", "unsort", "to_screen", "samples", "consts",
           "residuals", "KernelError", "NSPLATPARAMS", "NPOLARIZEDPARAMS", "NPOWERLAWPARAMS", "NKAPPAPARAMS",
           "gravitational_radius", "grid_axis", "library_path", "gradient_forward", "gradient_reverse", "vis_chi2_seed",
           "image_chi2_seed", "fit", "StokesMovie", "FrameScans", "chi2", "chi2_gradient", "sweep_passes", "frame_chi2_seed",
           "image_loss_gradient", "vis_seed", "closure_phases", "log_closure_amplitudes", "ClosureData", "closure_chi2",
           "frame_chi2_closures", "gain_factors", "apply_gains", "tiles", "Binning", "bin", "binned_grid", "binned_polar", "concatenate", "spacetime_tangent", "spacetime_tangent_differences", "spacetime_valgrad", "spacetime_jacobian",
           "spacetime_movie_loss", "ScatteringKernel", "taper", "RayLists", "ray_lists", "framed", "flattened", "frame_lists", "concatenate_screens", "screen_tiles", "tuning", "autotune", "queue_costs", "FitGraph", "march_precision"]
