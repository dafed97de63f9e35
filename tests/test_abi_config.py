"""CPU checks of the boundary: liblbm.so loads and exports every symbol of
include/lbm.h, the ctypes mirror of lbm_config_t / lbm_counters_t matches
the C layout, and the LBM_* parser restates the reference's derived
constants (LBM_3D_LES_air_over_waves_CODEX.py:35-529)."""
import ctypes
import math
import re
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import lbm_abi as abi
from lbm_config import parse_env, preset_env
from lbm_native import EXPORTS, LIB_PATH, load_library

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "lbm.h"


def header_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(lbm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    if not LIB_PATH.exists():
        pytest.skip("liblbm.so not built")
    lib = load_library()
    declared = header_functions()
    assert set(declared) == set(EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}\b", out), name


def test_struct_layout_matches_c():
    fields = [n for n, _ in abi.LbmConfig._fields_]
    cnt = [n for n, _ in abi.LbmCounters._fields_]
    ini = [n for n, _ in abi.LbmInitParams._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    src.append('printf("%zu %zu %zu\\n", sizeof(lbm_config_t), sizeof(lbm_counters_t), sizeof(lbm_init_params_t));')
    for n in fields:
        src.append(f'printf("c {n} %zu\\n", offsetof(lbm_config_t, {n}));')
    for n in cnt:
        src.append(f'printf("k {n} %zu\\n", offsetof(lbm_counters_t, {n}));')
    for n in ini:
        src.append(f'printf("i {n} %zu\\n", offsetof(lbm_init_params_t, {n}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "l.c"
        c.write_text("\n".join(src))
        exe = Path(d) / "l"
        subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
        lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    a, b, i = map(int, lines[0].split())
    assert a == ctypes.sizeof(abi.LbmConfig)
    assert b == ctypes.sizeof(abi.LbmCounters)
    assert i == ctypes.sizeof(abi.LbmInitParams)
    for line in lines[1:]:
        if not line:
            continue
        kind, name, off = line.split()
        struct = {"c": abi.LbmConfig, "k": abi.LbmCounters, "i": abi.LbmInitParams}[kind]
        assert getattr(struct, name).offset == int(off), name


def test_field_shapes():
    assert abi.field_shape(abi.FIELD_F, 8, 4, 6) == (19, 6, 4, 8)
    assert abi.field_shape(abi.FIELD_WATER_H, 8, 4, 6) == (4, 8)
    assert abi.field_shape(abi.FIELD_INFLOW_UX, 8, 4, 6) == (6, 4)
    assert abi.field_shape(abi.FIELD_RHO, 8, 4, 6) == (6, 4, 8)


def test_reference_defaults_and_derived_constants():
    """Default air run: CODEX:35-529 closed forms."""
    cfg = parse_env({})
    assert (cfg.Nx, cfg.Ny, cfg.Nz) == (600, 20, 40)
    wave_speed = math.sqrt(9.81 * 10.0)
    assert cfg.wave_speed == pytest.approx(wave_speed)
    assert cfg.U_ref_phys == 30.0
    assert cfg.vel_scale == pytest.approx(30.0 / 0.075)
    assert cfg.dt_phys == pytest.approx(0.075 * 0.5 / 30.0)
    nu0 = 0.075 * 40 / (30.0 * 20.0 / 1.516e-5)
    assert cfg.nu0 == pytest.approx(nu0)
    assert cfg.tau0 == pytest.approx(0.5 + 3 * nu0)
    assert cfg.offset_m == pytest.approx(2 * 1.5 + 3 * 0.5)
    assert cfg.collision_model == "kbc" and cfg.use_LES and cfg.C_backscatter == 0.02
    assert cfg.enable_top_drive == 1 and cfg.enable_x_fringe == 1 and cfg.enable_top_sponge == 1
    assert cfg.Cs == 0.2


def test_water_defaults():
    cfg = parse_env({"LBM_PHYSICS_MODE": "water", "LBM_FREE_SURFACE_TRACKING": "vof", "LBM_VOF_INITIAL_SHAPE": "block"})
    assert cfg.pressure_formulation == "vof_component_balanced"
    assert cfg.collision_model == "regularized" and not cfg.use_LES
    g = 9.81
    assert cfg.g_lb == pytest.approx(g * cfg.dt_phys**2 / cfg.dx)


@pytest.mark.parametrize(
    "env,msg",
    [
        ({"LBM_PHYSICS_MODE": "fire"}, "LBM_PHYSICS_MODE must be one of"),
        ({"LBM_COLLISION": "bgk"}, "LBM_COLLISION must be 'kbc' or 'regularized'"),
        ({"LBM_VOF_ORPHAN_SEARCH_RADIUS": "7"}, "LBM_VOF_ORPHAN_SEARCH_RADIUS must be in [1, 4]"),
        ({"LBM_VOF_INITIAL_SHAPE": "block"}, "LBM_VOF_INITIAL_SHAPE requires LBM_FREE_SURFACE_TRACKING=vof"),
        ({"LBM_TAU": "0.4"}, "LBM_TAU must be greater than 0.5"),
    ],
)
def test_validation_errors(env, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        parse_env(env)


def test_presets():
    c1 = parse_env(preset_env("plain_64x32x32"))
    assert (c1.Nx, c1.Ny, c1.Nz) == (64, 32, 32)
    assert c1.tau0 == 0.52 and c1.Cs == 0.16 and c1.collision_model == "regularized"
    c2 = parse_env(preset_env("waves_256x128x128"))
    assert (c2.Nx, c2.Ny, c2.Nz) == (256, 128, 128)
    assert c2.bed_phase_speed_mps / c2.vel_scale == pytest.approx(0.02)
    c4 = parse_env(preset_env("vof_192x96x160"))
    assert (c4.Nx, c4.Ny, c4.Nz) == (192, 96, 160) and c4.g_lb == 1e-5
    assert c4.free_surface_base_m == 80 and c4.vof_block_center_z_m == 128


def test_partial_flood_fill_is_refused():
    """A flood-fill sweep count below Nx + Ny + Nz would cut the fill short
    (schedule-dependent in the reference); it is refused, larger counts are
    accepted and passed through as the bound on the sweeps."""
    import pytest

    from harness import env_of
    from lbm_config import parse_env

    env = env_of("vof_192x96x160", LBM_LX_M=32, LBM_LY_M=16, LBM_LZ_M=40)
    with pytest.raises(ValueError, match="FLOOD_SWEEPS"):
        parse_env(dict(env, LBM_VOF_COLLAPSE_FLOOD_SWEEPS="5"))
    assert parse_env(dict(env, LBM_VOF_COLLAPSE_FLOOD_SWEEPS="1000")).to_struct().vof_collapse_flood_sweeps == 1000
    assert parse_env(env).to_struct().vof_collapse_flood_sweeps == 32 + 16 + 40
