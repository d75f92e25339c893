"""GPU parity of the C++ host job mirror (woodenpbrengine_amd/host/wpbr_jobs.hpp):
the reference's JOB_RUN sequences MEngine::buildLBVH and runCollisionSystem
run through the staged C ABI by a plain C++ program, checked against the
oracle: node arrays and shapesEntities bit-exact, interactions bit-exact
(hit shape and t; p = o + d t and wo = -d in float32)."""
import subprocess

import numpy as np
import pytest

import pyoracle as po
from woodenpbrengine_amd import PACKAGE_DIR, scenes

pytestmark = pytest.mark.gpu

DEMO = PACKAGE_DIR / "host" / "wpbr_jobs_demo"


def test_job_sequences_match_oracle(tmp_path):
    assert DEMO.exists(), "build the host demo first (__graft_entry__.build())"
    tris = scenes.random_triangles(30000, 41, median_edge=0.01)
    g = np.random.default_rng(42)
    sx = np.stack([scenes.translate_xform(c) for c in g.random((3, 3))]).astype(np.float32)
    sr = np.array([0.05, 0.08, 0.03], np.float32)
    rays = np.concatenate([scenes.pinhole_rays(64, 32, (0.5, 0.5, -1.0)),
                           scenes.random_rays(4096, 43)], axis=1).astype(np.float32)
    n = rays.shape[1]
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([len(sr), len(tris), n], np.uint32).tobytes())
        f.write(sx.astype(np.float32).tobytes())
        f.write(sr.tobytes())
        f.write(tris.astype(np.float32).tobytes())
        f.write(np.ascontiguousarray(rays.T).tobytes())
    out = tmp_path / "out.bin"
    print("running", DEMO, flush=True)
    subprocess.run([str(DEMO), str(inp), str(out)], check=True, timeout=120)
    buf = out.read_bytes()
    nn, ns = np.frombuffer(buf, np.uint32, 2)
    o = 8
    nodes = np.frombuffer(buf, np.uint32, 3 * nn, o).reshape(nn, 3); o += 12 * nn
    bounds = np.frombuffer(buf, np.float32, 6 * nn, o).reshape(nn, 6); o += 24 * nn
    ents = np.frombuffer(buf, np.uint32, ns, o); o += 4 * ns
    inter = np.frombuffer(buf, np.float32, 8 * n, o).reshape(n, 8)

    osc = po.OScene(sx, sr, tris)
    ob = po.build_scene_bvh(osc)
    assert nn == ob.n_nodes
    np.testing.assert_array_equal(nodes[:, 0], ob.get("node_offset"))
    np.testing.assert_array_equal(nodes[:, 1], ob.get("node_nshapes"))
    interior = nodes[:, 1] == 0
    np.testing.assert_array_equal(nodes[interior, 2], ob.get("node_axis")[interior])
    np.testing.assert_array_equal(bounds, ob.get("node_bounds"))
    np.testing.assert_array_equal(ents, ob.get("sorted_index"))

    os_, ot, _, _ = po.collide(osc, ob, rays)
    hit = inter[:, 0].view(np.uint32)
    np.testing.assert_array_equal(hit, os_)
    np.testing.assert_array_equal(inter[:, 1].view(np.uint32), ot.view(np.uint32))
    m = hit != 0xFFFFFFFF
    p = rays[:3].T[m] + rays[3:].T[m] * ot[m, None]
    np.testing.assert_array_equal(inter[m, 2:5], p.astype(np.float32))
    np.testing.assert_array_equal(inter[m, 5:8], -rays[3:].T[m])
