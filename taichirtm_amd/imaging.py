"""Imaging utilities: load, stack and prepare RTM images (src/imaging.py:29-277).

Host-side NumPy post-processing with the reference's semantics.  Stacking sums
per-shot images in float32 in shot order, transposes to (nz, nx) and removes
the global mean (imaging.py:57-129); ``stack_images_device`` is the on-device
form for images that are still resident in HBM.
"""
from __future__ import annotations

import glob
import os
from typing import Any, Dict, List, Optional

import numpy as np


def load_rtm_results(directory: str, pattern: str = "*.npz") -> List[Dict[str, Any]]:
    """Load RTM result files (imaging.py:29-54).  Files are read without
    unpickling; an entry stored as a pickled object (the reference saves
    ``surface_matrix=None`` that way) is returned as None."""
    results = []
    for filepath in glob.glob(os.path.join(directory, pattern)):
        with np.load(filepath, allow_pickle=False) as data:
            result = {}
            for key in data.files:
                try:
                    result[key] = data[key]
                except ValueError:
                    result[key] = None
        result["filepath"] = filepath
        results.append(result)
    return results


def stack_rtm_images(results: List[Dict[str, Any]], subtract_mean: bool = True) -> Dict[str, Any]:
    """Stack per-shot images (imaging.py:57-129)."""
    if not results:
        raise ValueError("No results to stack")
    u_sum = v_sum = w_sum = None
    for i, data in enumerate(results):
        u, v, w = data["u"], data["v"], data["w"]
        if i == 0:
            u_sum, v_sum, w_sum = np.zeros_like(u), np.zeros_like(v), np.zeros_like(w)
        u_sum += u
        v_sum += v
        w_sum += w
    umap, vmap, wmap = u_sum.T, v_sum.T, w_sum.T
    if subtract_mean:
        umap = umap - np.mean(umap)
        vmap = vmap - np.mean(vmap)
        wmap = wmap - np.mean(wmap)
    last = results[-1]
    offset, dx, dz = float(last["offset"]), float(last["dx"]), float(last["dz"])
    nx, nz = int(last["nx"]), int(last["nz"])
    xmin, xmax, zmin, zmax = -offset, dx * nx - offset, 0.0, dz * nz
    return {"u": umap, "v": vmap, "w": wmap, "xmin": xmin, "xmax": xmax, "zmin": zmin,
            "zmax": zmax, "extent": [xmin, xmax, zmax, zmin], "dx": dx, "dz": dz,
            "nx": nx, "nz": nz}


def stack_images_device(images, dtype="float64"):
    """Sum a sequence of (nx, nz) device images (torch tensors) in the given
    precision, in order; returns the device sum (before transpose / mean)."""
    import torch
    acc = None
    for img in images:
        x = img.to(getattr(torch, dtype))
        acc = x.clone() if acc is None else acc.add_(x)
    return acc


def compute_display_limits(data: np.ndarray) -> tuple:
    """Symmetric colour limits (imaging.py:132-147)."""
    data_max = np.max(np.abs(data))
    return -data_max, data_max


def prepare_image_data(image: np.ndarray, attenuate_region: Optional[tuple] = None,
                       attenuation_factor: float = 1e-3) -> np.ndarray:
    """Copy of ``image`` with a clipped region scaled (imaging.py:150-182)."""
    result = image.copy()
    if attenuate_region is not None:
        x0, x1, z0, z1 = attenuate_region
        x0, x1 = max(0, x0), min(result.shape[0], x1)
        z0, z1 = max(0, z0), min(result.shape[1], z1)
        if x0 < x1 and z0 < z1:
            result[x0:x1, z0:z1] *= attenuation_factor
    return result


def create_visualization_data(rtm_instance, subtract_mean: bool = True,
                              attenuate_source: bool = True,
                              source_attenuation_radius: int = 10) -> Dict[str, Any]:
    """Display-ready images from a finished ReverseTimeMigration (imaging.py:185-248)."""
    u = rtm_instance.image_u.copy()
    v = rtm_instance.image_v.copy()
    w = rtm_instance.image_w.copy()
    if attenuate_source and hasattr(rtm_instance, "src_loc_step"):
        src = rtm_instance.src_loc_step[0]
        r = source_attenuation_radius
        x0, x1 = max(0, src[0] - r), min(u.shape[0], src[0] + r)
        z0, z1 = src[1], min(u.shape[1], src[1] + r)
        u[x0:x1, z0:z1] *= 1e-3
        v[:, z0:z1] *= 1e-1
        w[x0:x1, z0:z1] *= 1e-3
    umap, vmap, wmap = u.T, v.T, w.T
    if subtract_mean:
        umap = umap - np.mean(umap)
        vmap = vmap - np.mean(vmap)
        wmap = wmap - np.mean(wmap)
    ext = rtm_instance.get_axes_extent()
    return {"u": umap, "v": vmap, "w": wmap, "xmin": ext["xmin"], "xmax": ext["xmax"],
            "zmin": ext["zmin"], "zmax": ext["zmax"],
            "extent": [ext["xmin"], ext["xmax"], ext["zmax"], ext["zmin"]],
            "u_lim": compute_display_limits(umap), "v_lim": compute_display_limits(vmap),
            "w_lim": compute_display_limits(wmap)}


def save_stacked_results(data: Dict[str, Any], filepath: str):
    """Write a stacked image npz (imaging.py:251-277)."""
    directory = os.path.dirname(filepath)
    if directory and not os.path.exists(directory):
        os.makedirs(directory)
    np.savez_compressed(filepath, u=data["u"], v=data["v"], w=data["w"], xmin=data["xmin"],
                        xmax=data["xmax"], zmin=data["zmin"], zmax=data["zmax"])
    print(f"Stacked results saved to {filepath}")
