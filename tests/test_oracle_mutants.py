"""The material pins bite: rt_oracle.c rebuilt with one planted mistake at a time must fail
tests/test_oracle_materials.py, and the same recipe without a mistake must pass it.

Each mistake is a plausible slip in the METALLIC / DIELECTRIC scatter (the A5 extension of
raytraceBVH.comp:100-115).  The mistakes are planted in a copy of the oracle's source in a temporary
directory, compiled with the oracle's own flags, and loaded by a child pytest through RT_ORACLE_LIB;
nothing under oracle/ is modified.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle" / "rt_oracle.c"
CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared"]

MUTANTS = {
    "refraction ratio with the face flag inverted":
        ("float ri = backface ? ior : (1.0f / ior);", "float ri = backface ? (1.0f / ior) : ior;"),
    "fuzz added after the normalisation":
        ("nd = normalize3(add3(refl, smul(m->albedo[3], rs)));", "nd = add3(normalize3(refl), smul(m->albedo[3], rs));"),
    "Schlick fed the wrong-signed cosine":
        ("float ct = gmin(dot3(neg3(d), n), 1.0f);", "float ct = gmin(dot3(d, n), 1.0f);"),
    "metal reflecting the reversed ray":
        ("v3 refl = reflect3(d, n);", "v3 refl = reflect3(neg3(d), n);"),
    "total internal reflection drawing a number":
        ("if (cannot || schlick(ct, ri) > orc_rng_next(st))", "if (schlick(ct, ri) > orc_rng_next(st) || cannot)"),
}


def _build(tmp: Path, find: str | None, repl: str | None) -> Path:
    text = SRC.read_text()
    if find is not None:
        assert text.count(find) == 1, f"the mutation site is not unique in rt_oracle.c: {find}"
        text = text.replace(find, repl)
    src = tmp / "rt_oracle.c"
    src.write_text(text)
    lib = tmp / "liboracle_variant.so"
    subprocess.run(["gcc", *CFLAGS, f"-I{ROOT / 'oracle'}", "-o", str(lib), str(src), "-lm"], check=True,
                   capture_output=True)
    return lib


def _run_material_pins(lib: Path) -> subprocess.CompletedProcess:
    env = dict(os.environ, RT_ORACLE_LIB=str(lib))
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                           str(ROOT / "tests" / "test_oracle_materials.py")],
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)


def test_unmutated_rebuild_passes(tmp_path):
    r = _run_material_pins(_build(tmp_path, None, None))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_fails_the_material_pins(tmp_path, name):
    find, repl = MUTANTS[name]
    r = _run_material_pins(_build(tmp_path, find, repl))
    # exit code 1: the tests ran and some failed (not an import or collection error)
    assert r.returncode == 1, (name, r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert " failed" in r.stdout, r.stdout[-2000:]
