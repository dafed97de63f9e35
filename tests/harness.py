"""Shared setup for parity tests: one initial state for the device and the
oracle (both start from the same uploaded macroscopic fields)."""
from __future__ import annotations

import numpy as np

import lbm_abi as abi
from lbm_config import parse_env, preset_env
from lbm_init import initial_fields
from oracle_ref import OracleCase


def env_of(preset: str, **overrides) -> dict:
    env = preset_env(preset)
    env.update({k: str(v) for k, v in overrides.items()})
    return env


def setup_case(case, cfg_values, piecewise=None):
    """Run the init sequence on a device Case or an OracleCase."""
    if cfg_values.bed_profile == "piecewise_linear":
        case.upload(abi.FIELD_PIECEWISE_BED_Z, cfg_values.piecewise_bed_z_np)
    case.geometry(0.0)
    obstacle0 = case.download(abi.FIELD_OBSTACLE)
    fields = initial_fields(cfg_values, obstacle0, case.download(abi.FIELD_WATER_H))
    for fid, arr in fields.items():
        case.upload(fid, arr)
    case.init_populations()
    return fields


def make_oracle(env: dict, fp64: bool = False):
    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct(), fp64=fp64)
    setup_case(o, cfg)
    return cfg, o
