"""The oracle pins catch plausible mistakes: each mistake below is planted in a
scratch copy of the C++ oracle (built with its own Makefile), and the pins of
tests/test_oracle_pins.py run against that copy must fail.  CPU only; the
repository's oracle is not touched.
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

# name: (line of oracle/cmf_oracle.cpp, the same line with the mistake)
MISTAKES = {
    "friction exponent 4/3 instead of 7/3": (
        "* ((REAL)1.0 / (r_sifd * r_sifd * cbrt_real(r_sifd)));",
        "* ((REAL)1.0 / (r_sifd * cbrt_real(r_sifd)));"),
    "backflow limit 10 % instead of 5 %": (
        "const REAL limit = fmin_(negative ? (REAL)0.05 * total_storage / total_neg : (REAL)1.0, (REAL)1.0);",
        "const REAL limit = fmin_(negative ? (REAL)0.10 * total_storage / total_neg : (REAL)1.0, (REAL)1.0);"),
    "storage limiter inverted": (
        "lim[i] = (og > (REAL)1e-8) ? fmin_(rate / og, (REAL)1.0) : (REAL)1.0;",
        "lim[i] = (og > (REAL)1e-8) ? fmin_(og / rate, (REAL)1.0) : (REAL)1.0;"),
    "path net added at both ends": (
        "gbif[di] += (STO)(-net);",
        "gbif[di] += (STO)(net);"),
    "river deficit not carried into flood storage": (
        "STO fsu = fs + ((rsu < (STO)0.0) ? rsu : (STO)0.0) + (STO)((fin - fout - gbif) * time_step);",
        "STO fsu = fs + (STO)((fin - fout - gbif) * time_step);"),
    "floodplain width index off by one": (
        "const REAL W = rw + (REAL)(k + 1) * width_increment;",
        "const REAL W = rw + (REAL)(k) * width_increment;"),
    "CFL offset sign": (
        "const REAL nf = (REAL)std::floor(outer_time_step / dtc - (REAL)0.01) + (REAL)1.0;",
        "const REAL nf = (REAL)std::floor(outer_time_step / dtc + (REAL)0.01) + (REAL)1.0;"),
    "CFL depth floor 0.1 m": (
        "const REAL depth = fmax_(river_depth[i], (REAL)0.01);",
        "const REAL depth = fmax_(river_depth[i], (REAL)0.1);"),
    "reservoir conservation release without the square root": (
        "out = qn * (REAL)std::sqrt(total / cv);",
        "out = qn * (total / cv);"),
    "path limit on the larger end storage": (
        "const REAL limit = fmin_((REAL)0.05 * fmin_(ts0, ts1) / (std::fabs(sum) * time_step), (REAL)1.0);",
        "const REAL limit = fmin_((REAL)0.05 * fmax_(ts0, ts1) / (std::fabs(sum) * time_step), (REAL)1.0);"),
    "levee case 2 without the channel width": (
        "const REAL dwth_add_c2 = levee_distance + rw;",
        "const REAL dwth_add_c2 = levee_distance;"),
    "protected-side threshold sign": (
        "const REAL s_top = base_sto + (levee_distance + rw) * dhgt_dif * rl;",
        "const REAL s_top = base_sto - (levee_distance + rw) * dhgt_dif * rl;"),
    "levee case 4 measured from the base": (
        "const REAL dsto_add_c4 = (f_dph_c4 - crown) * (levee_distance + rw) * rl;",
        "const REAL dsto_add_c4 = (f_dph_c4 - base) * (levee_distance + rw) * rl;"),
    "precise cross-section depth measured from the bank": (
        "upd_rcsd = mouth ? r_dep : (REAL)(mw - rel);",
        "upd_rcsd = mouth ? r_dep : (REAL)(mw - (double)c_elv);"),
    "precise surface without the river height": (
        "const double rel = (double)c_elv - (double)r_hgt;",
        "const double rel = (double)c_elv;"),
    "levee fill scan on the base height": (
        "const REAL ratio_fill = (crown - dhgt_pre) / dhgt_seg;",
        "const REAL ratio_fill = (base - dhgt_pre) / dhgt_seg;"),
    "levee path levels above the channel on the river surface": (
        "upd_csd = fmax_(precise ? (REAL)(max_pwd - (double)bel[k]) : max_pw - bel[k], (REAL)0.0);",
        "upd_csd = fmax_(precise ? (REAL)(max_pwd - (double)bel[k]) : max_w - bel[k], (REAL)0.0);"),
    "levee path channel level falls back to the 1 cm floor": (
        "sifd = (semi <= (REAL)0.0) ? upd_csd : semi;",
        "sifd = (semi <= (REAL)0.0) ? (REAL)std::sqrt(upd_csd * (REAL)0.01) : semi;"),
    "levee path levels above the channel semi-implicit": (
        "sifd = upd_csd;",
        "sifd = (REAL)std::sqrt(upd_csd * csd_old);"),
    "kinematic flow with depth to the 4/3": (
        "const REAL kin_riv_vel = ((REAL)1.0 / r_man) * (REAL)std::sqrt(bed_slope) * cbrt_real(r_dep * r_dep);",
        "const REAL kin_riv_vel = ((REAL)1.0 / r_man) * (REAL)std::sqrt(bed_slope) * cbrt_real(r_dep);"),
    "kinematic floodplain slope not capped": (
        "const REAL bed_slope_f = fmin_(bed_slope, (REAL)0.005);",
        "const REAL bed_slope_f = bed_slope;"),
    "kinematic bed slope not floored": (
        "bed_slope = fmax_(bed_slope, min_kin);",
        "bed_slope = fabs(bed_slope);"),
    "log inflow sum without the flood inflow": (
        "q[L_INFLOW] = (double)((rin + fin + inflow_extra) * time_step);",
        "q[L_INFLOW] = (double)((rin + inflow_extra) * time_step);"),
    "log stage sums include levee catchments": (
        "if (non_levee) {",
        "if (non_levee || true) {"),
    "log slot overwritten instead of accumulated": (
        "sums[k * stride + current_step] += acc * 1e-9;",
        "sums[k * stride + current_step] = acc * 1e-9;"),
    "log inflow error with the sign reversed": (
        "q[L_INFLOW_ERROR] = (double)(total_pre - total_next",
        "q[L_INFLOW_ERROR] = (double)(total_next - total_pre"),
    "path levels rescaled by the opposite ends": (
        "bout[k] = (raw >= (REAL)0.0) ? raw * l0 : raw * l1;",
        "bout[k] = (raw >= (REAL)0.0) ? raw * l1 : raw * l0;"),
    "path net always scaled by the upstream end": (
        "const REAL net = (raw_sum >= (REAL)0.0) ? raw_sum * l0 : raw_sum * l1;",
        "const REAL net = raw_sum * l0;"),
}


@pytest.mark.parametrize("name", list(MISTAKES))
def test_pins_catch_a_planted_mistake(name, tmp_path):
    if shutil.which("make") is None or shutil.which(os.environ.get("CXX", "g++")) is None:
        pytest.skip("no C++ toolchain")
    old, new = MISTAKES[name]
    src = (ROOT / "oracle" / "cmf_oracle.cpp").read_text()
    assert src.count(old) == 1, "the line to alter is no longer in the oracle"
    work = tmp_path / "copy"
    shutil.copytree(ROOT / "oracle", work / "oracle", ignore=shutil.ignore_patterns("*.so", "__pycache__"))
    (work / "oracle" / "cmf_oracle.cpp").write_text(src.replace(old, new))
    subprocess.run(["make", "-s", "-C", str(work / "oracle")], check=True)
    (work / "tests").mkdir()
    for f in ("conftest.py", "test_oracle_pins.py"):
        shutil.copy(ROOT / "tests" / f, work / "tests" / f)
    # one thread: the pins are tiny, and parallel regions would only spin
    env = dict(os.environ, PYTHONPATH=f"{work}{os.pathsep}{ROOT / 'cama-flood-gpu_amd'}", OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        str(work / "tests" / "test_oracle_pins.py")],
                       capture_output=True, text=True, env=env, cwd=work)
    assert r.returncode == 1 and " failed" in r.stdout, r.stdout[-2000:]
