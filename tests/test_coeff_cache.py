"""The reference's on-disk matrix cache (src/elastodynamics_cqbem_matrix_collection.cpp:41-57,77-94):
folder <coeffs_folder_path>/elastodynamicsCQBEM_<key>, key = upper-case hex std::hash of the object group's
compact JSON dump without simulation.compression_ratio / update_local_frame; G_n.mat / H_n.mat in the
include/matrix_io.hpp format.

Golden vectors: the keys and dumps of the reference's four shipped scenes (tests/golden/scenes/, copied from
its data/), written by oracle/_ref/config_key -- a checker compiled against the reference's own vendored
JSON library (tests/golden/make_scene_cache_keys.sh).  Where that checker is built, the number formatting is
also compared on thousands of random doubles."""
import json
import os
import random
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DRIVER = os.path.join(ROOT, "surfaceonlydeformables_amd", "collection_driver")
CHECKER = os.path.join(ROOT, "oracle", "_ref", "config_key")


def driver(*args):
    r = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_shipped_scene_cache_keys():
    rows = [l.split("\t") for l in open(os.path.join(GOLDEN, "scene_cache_keys.txt")).read().splitlines()]
    assert len(rows) == 4
    for scene, key, dump in rows:
        got = driver("--cache-key", os.path.join(GOLDEN, "scenes", scene)).splitlines()
        assert got == [key + "\t" + dump], scene


@pytest.mark.skipif(not os.path.exists(CHECKER), reason="oracle/_ref/config_key not built (needs the reference tree)")
def test_json_number_format_matches_reference_library(tmp_path):
    """Doubles of every magnitude (and random bit patterns) print digit for digit as the reference's JSON
    library prints them (Grisu2 digits, fixed/exponent layout, '.0' on integral values), plus integers,
    escapes and nesting."""
    rng = random.Random(7)
    vals = [rng.choice([1, -1]) * 10 ** rng.uniform(-300, 300) * rng.random() for _ in range(4000)]
    while len(vals) < 8000:
        v = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
        if np.isfinite(v):
            vals.append(v)
    vals += [0.0, -0.0, 1.0, 0.1, 1e-5, 1e-4, 1e15, 1e16, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
             123456789012345.0, 1234567890123456.0, 0.004999999888241291, 0.44999998807907104]
    doc = {"f": vals, "i": [0, -1, 7, 123456789012345678, -9223372036854775807],
           "s": ["a\"b\\c\n\t\u0001/é\U0001F600"], "b": [True, False, None], "o": {"z": 1.5, "a": {}, "m": []}}
    path = tmp_path / "doc.json"
    path.write_text(json.dumps(doc))
    ref = subprocess.run([CHECKER, "--dump", str(path)], capture_output=True, text=True, timeout=60).stdout
    assert driver("--dump", str(path)) == ref


@pytest.mark.skipif(not os.path.exists(CHECKER), reason="oracle/_ref/config_key not built (needs the reference tree)")
def test_cache_key_of_edited_scene(tmp_path):
    """Keys follow every edit the reference's hash sees, and ignore the two keys it erases."""
    base = json.load(open(os.path.join(GOLDEN, "scenes", "cube_drop.json")))
    variants = []
    for edit in range(6):
        sc = json.loads(json.dumps(base))
        sim = sc["physics_object_groups"][0]["physics_object_group"]["simulation"]
        if edit == 1:
            sim["dt"] = 0.001
        elif edit == 2:
            sim["compression_ratio"] = 0.25  # erased before hashing
        elif edit == 3:
            sim["update_local_frame"] = False  # erased before hashing
        elif edit == 4:
            sim["gaussian_quadrature_order"] = 3
        elif edit == 5:
            sc["physics_object_groups"][0]["physics_object_group"]["material"]["nu"] = 0.3
        path = tmp_path / f"s{edit}.json"
        path.write_text(json.dumps(sc, indent=2))
        ref = subprocess.run([CHECKER, str(path)], capture_output=True, text=True, timeout=60).stdout
        got = driver("--cache-key", str(path))
        assert got == ref, edit
        variants.append(got.split("\t")[0])
    assert variants[0] == variants[2] == variants[3]
    assert len({variants[0], variants[1], variants[4], variants[5]}) == 4
