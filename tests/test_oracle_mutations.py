"""The oracle's known-answer pins catch planted mistakes.

Each case below plants one plausible mistake in a temporary copy of
oracle/oracle.cpp (a swapped blend weight, a negated or ungated aux choice, a
wrong comparison, a misplaced timewarp plane, a dropped guard), builds it
with g++, points the oracle binding at the broken library
(ISTI_ORACLE_LIB) and runs the pins in a child process: every mutant must
fail at least one of them, and the unmodified source must pass them all.
Mistakes are planted only in the CPU oracle, never in GPU code.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle" / "oracle.cpp"
PINS = ["tests/test_oracle_pins.py", "tests/test_oracle_kat.py"]
FLAGS = ["-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared",
         f"-I{ROOT / 'oracle'}"]

# (name, text in oracle.cpp, replacement)
MUTANTS = [
    ("blend weights swapped",
     "r.color = second_chance(1.0f - alpha, frame_idx, sidx, st, fi, s, fwd, bwd, eps2);\n        } else {",
     "r.color = second_chance(alpha, frame_idx, sidx, st, fi, s, fwd, bwd, eps2);\n        } else {"),
    ("bwd blend weight 1-alpha",
     "r.color = second_chance(alpha, frame_idx, sidx, st, bi, s, bwd, fwd, eps2);",
     "r.color = second_chance(1.0f - alpha, frame_idx, sidx, st, bi, s, bwd, fwd, eps2);"),
    ("lerp operands swapped",
     "out = lerp_to_u8(color_from_shadow(sidx, st, s, to, up), out, weight);",
     "out = lerp_to_u8(out, color_from_shadow(sidx, st, s, to, up), weight);"),
    ("choose_aux negated", "if (choose_aux) out[", "if (!choose_aux) out["),
    ("visibility gate dropped", "if (main_r.valid && !best.visible_in_main) {", "if (main_r.valid) {"),
    ("aux-vs-main similarity ignored", "choose_aux = !similar && best.depth < main_r.depth;",
     "choose_aux = best.depth < main_r.depth;"),
    ("aux priority strict", "ar.priority >= best.priority", "ar.priority > best.priority"),
    ("aux error non-strict", "ar.error < best.error", "ar.error <= best.error"),
    ("aux closer flipped", "(!similar && ar.depth < best.depth)", "(!similar && ar.depth > best.depth)"),
    ("timewarp plane through the mean-z point", "const V3 pp{campos.x, campos.y, campos.z};",
     "const V3 pp{center_ws.x, center_ws.y, center_ws.z};"),
    ("timewarp plane through the ray origin", "const V3 pp{campos.x, campos.y, campos.z};",
     "const V3 pp{ro.x, ro.y, ro.z};"),
    ("timewarp t negated", "const float t = dot3(", "const float t = -dot3("),
    ("timewarp parallel guard dropped", "if (std::fabs(denom) < 1e-6f) return;", "(void)denom;"),
    ("timewarp never falls back to aux", "if (in_range(pix, tw.res_to) || !aux.isset) {", "if (true) {"),
    ("timewarp aux not clamped to its last column",
     "I2{(int32_t)aux.res_to.x - 1, (int32_t)aux.res_to.y - 1}", "I2{(int32_t)aux.res_to.x - 2, (int32_t)aux.res_to.y - 1}"),
]


def _build(name: str, text: str, repl: str | None, out_dir: Path) -> Path:
    src = SRC.read_text()
    if repl is not None:
        assert src.count(text) == 1, f"mutation site of {name!r} not found exactly once"
        src = src.replace(text, repl)
    cpp = out_dir / f"{abs(hash(name))}.cpp"
    cpp.write_text(src)
    so = cpp.with_suffix(".so")
    res = subprocess.run(["g++", *FLAGS, str(cpp), "-o", str(so)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return so


def _run_pins(so: Path) -> subprocess.CompletedProcess:
    env = dict(os.environ, ISTI_ORACLE_LIB=str(so))
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *PINS],
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("mutants")
    jobs = [("original", "", None)] + MUTANTS
    with ThreadPoolExecutor(max_workers=4) as pool:
        sos = list(pool.map(lambda j: _build(*j, d), jobs))
    return dict(zip([j[0] for j in jobs], sos))


def test_unmodified_oracle_passes_the_pins(builds):
    res = _run_pins(builds["original"])
    assert res.returncode == 0, res.stdout[-2000:]


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_planted_mistake_fails_a_pin(builds, name):
    res = _run_pins(builds[name])
    assert res.returncode == 1, f"mutant {name!r} was not caught:\n{res.stdout[-1500:]}"
