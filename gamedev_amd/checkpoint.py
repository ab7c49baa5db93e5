"""Checkpoint files in the reference's formats, so that one initial state can feed both builds and their
outputs can be read with the same readers.

particle_<frame>.dat   N_P raw 64-byte records, no header (save_host_binary / load_host_binary,
                       inc/swarm/swarm_host.cuh:778-815): position x, y, z, velocity x, y, z, par_size,
                       par_numr as little-endian float64 ([SWARM_DTYPE], swarm_host.cuh:1124-1140). In the
                       file the azimuthal and polar velocities are linear; in memory they are angular
                       (save_sam_as_velocity / load_velocity_as_sam, swarm_host.cuh:866-902).
variables.txt          configparser text with [PARAMETERS] and [SWARM_DTYPE] (save_variable,
                       swarm_host.cuh:930-1141); the collision keys this build knows are written.
rngstate_<frame>.dat   this build's per-particle Philox draw counters (8 bytes each). The reference
                       stores hipRAND XORWOW states (40 bytes each) there; the two are not
                       interchangeable and a file of the wrong size is rejected.
Frame numbers are zero-padded to max(5, digits of SAVE_MAX) (frame_num, swarm_host.cuh:905-912).
"""
from __future__ import annotations

import configparser
import math
from pathlib import Path

import numpy as np

SWARM_FIELDS = ("position_x", "position_y", "position_z", "velocity_x", "velocity_y", "velocity_z",
                "par_size", "par_numr")
SWARM_DTYPE = np.dtype([(name, "<f8") for name in SWARM_FIELDS])
assert SWARM_DTYPE.itemsize == 64


def frame_num(idx_file: int, save_max: int = 100) -> str:
    width = max(5, len(str(save_max)))
    return str(idx_file).zfill(width)


def _cyl_r(y, z, n_z: int):
    return y*np.sin(z) if n_z > 1 else y


def to_file_records(x, y, z, vx, vy, vz, size, numr, info: dict) -> np.ndarray:
    """In-memory state (angular azimuthal / polar velocities) to file records (linear velocities)."""
    n = len(x)
    rec = np.empty(n, SWARM_DTYPE)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    rec["position_x"] = x if info["n_x"] > 1 else 0.5*(info["x_min"] + info["x_max"])
    rec["position_y"] = y
    rec["position_z"] = z if info["n_z"] > 1 else 0.5*math.pi
    R = _cyl_r(y, rec["position_z"], info["n_z"])
    with np.errstate(divide="ignore", invalid="ignore"):
        rec["velocity_x"] = np.where(R > 0.0, np.asarray(vx, np.float64)/R, 0.0)
        rec["velocity_z"] = (np.where(y > 0.0, np.asarray(vz, np.float64)/y, 0.0) if info["n_z"] > 1
                             else 0.0)
    rec["velocity_y"] = vy
    rec["par_size"] = size
    rec["par_numr"] = numr
    return rec


def from_file_records(rec: np.ndarray, info: dict) -> dict:
    """File records to the in-memory state (angular velocities), as load_velocity_as_sam."""
    x = np.array(rec["position_x"]) if info["n_x"] > 1 else np.full(len(rec), 0.5*(info["x_min"] + info["x_max"]))
    z = np.array(rec["position_z"]) if info["n_z"] > 1 else np.full(len(rec), 0.5*math.pi)
    y = np.array(rec["position_y"])
    R = _cyl_r(y, z, info["n_z"])
    vx = np.array(rec["velocity_x"])*R
    vz = np.array(rec["velocity_z"])*y if info["n_z"] > 1 else np.zeros(len(rec))
    return {"x": x, "y": y, "z": z, "vx": vx, "vy": np.array(rec["velocity_y"]), "vz": vz,
            "size": np.array(rec["par_size"]), "numr": np.array(rec["par_numr"])}


def save_particles(path, idx_file: int, state: dict, info: dict, save_max: int = 100) -> Path:
    """Write particle_<frame>.dat from a state dict (x, y, z, size, numr and optional vx, vy, vz)."""
    n = len(state["x"])
    zeros = np.zeros(n)
    rec = to_file_records(state["x"], state["y"], state["z"], state.get("vx", zeros), state.get("vy", zeros),
                          state.get("vz", zeros), state["size"], state["numr"], info)
    out = Path(path)/f"particle_{frame_num(idx_file, save_max)}.dat"
    rec.tofile(out)
    return out


def load_particles(path, idx_file: int, n: int, info: dict, save_max: int = 100) -> dict:
    """Read particle_<frame>.dat of exactly n records (a file of another size is an error)."""
    f = Path(path)/f"particle_{frame_num(idx_file, save_max)}.dat"
    if f.stat().st_size != n*SWARM_DTYPE.itemsize:
        raise ValueError(f"{f}: expected {n} records of {SWARM_DTYPE.itemsize} bytes")
    return from_file_records(np.fromfile(f, SWARM_DTYPE), info)


def save_rng(path, idx_file: int, counters: np.ndarray, save_max: int = 100) -> Path:
    out = Path(path)/f"rngstate_{frame_num(idx_file, save_max)}.dat"
    np.ascontiguousarray(counters, np.uint64).tofile(out)
    return out


def load_rng(path, idx_file: int, n: int, save_max: int = 100) -> np.ndarray:
    f = Path(path)/f"rngstate_{frame_num(idx_file, save_max)}.dat"
    if f.stat().st_size != 8*n:
        raise ValueError(f"{f}: expected {n} Philox draw counters of 8 bytes (hipRAND XORWOW states are "
                         "not readable by this build)")
    return np.fromfile(f, np.uint64)


def save_variables(path, info: dict, n: int, total_dust_mass: float, search: str = "morton") -> Path:
    """variables.txt with the collision parameters of the model and the swarm record layout."""
    cfg = configparser.ConfigParser()
    cfg.optionxform = str
    geometry = {(False, False): "radial", (True, False): "radial_azimuthal", (False, True): "radial_polar",
                (True, True): "full_3d"}[(info["n_x"] > 1, info["n_z"] > 1)]
    p = {
        "LAMBDA_0": f"{n/info['n_k']/total_dust_mass:.8e}", "V_FRAG": f"{info['v_frag']:.8e}",
        "COAG_KERNEL": str(info["coag_kernel"]), "N_K": str(info["n_k"]), "H_SEARCH": f"{info['h_search']:.8g}",
        "COL_BATH_MAX": f"{info['col_bath_max']:.8e}", "COL_BATH_EPS": f"{info['col_bath_eps']:.8e}",
        "COL_BATH_ALPHA": f"{info['col_bath_alpha']:.8e}", "COLLISION_SEARCH": search,
        "MORTON_LEAF_TARGET": str(info["morton_leaf_target"]), "MORTON_MAX_LEVEL": str(info["morton_max_level"]),
        "TOTAL_DUST_MASS": f"{total_dust_mass:.8e}", "RHO_0": f"{info['rho_0']:.8e}", "N_P": f"{float(n):.8e}",
        "N_X": str(info["n_x"]), "X_MIN": f"{info['x_min']:.8g}", "X_MAX": f"{info['x_max']:.8g}",
        "N_Y": str(info.get("n_y", 1)), "Y_MIN": f"{info['y_min']:.8g}", "Y_MAX": f"{info['y_max']:.8g}",
        "N_Z": str(info["n_z"]), "Z_MIN": f"{info['z_min']:.8g}", "Z_MAX": f"{info['z_max']:.8g}",
        "GEOMETRY": geometry, "DUST_DENSITY_KIND": "surface" if info["n_z"] == 1 else "volume",
        "INIT_SMIN": f"{info['init_smin']:.8e}", "RNG_STREAM_POLICY": "philox_counter_per_particle",
    }
    cfg["PARAMETERS"] = p
    cfg["SWARM_DTYPE"] = {name: "f8" for name in SWARM_FIELDS}
    out = Path(path)/"variables.txt"
    with open(out, "w") as fh:
        cfg.write(fh)
    return out


def read_swarm_dtype(path) -> np.dtype:
    """The record dtype a reader builds from variables.txt (as the reference's scripts do)."""
    cfg = configparser.ConfigParser()
    cfg.read(Path(path)/"variables.txt")
    return np.dtype([(name, value) for name, value in cfg["SWARM_DTYPE"].items()])
