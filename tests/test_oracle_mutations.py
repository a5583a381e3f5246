"""The builder KATs (test_oracle_kat_builders.py) are sharp: each of the
reference's builder rules is flipped in a copy of the oracle's source
(oracle/bvh.c), the copy is compiled into a temporary library, and the KAT
that names the rule must fail against it.  The shipped oracle is untouched.

Flips: `>` for `>=` termination; `<=` for `<` in the split search; axes or
planes scanned from the other end; a 1/3 centroid factor; a stable
partition; right-first and breadth-first child allocation; the TLAS walk
following the moved entry; `<=` in the TLAS match.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

FLIPS = {
    "ge_to_gt": ([("if (split_cost >= parent_cost) continue;", "if (split_cost > parent_cost) continue;")],
                 "test_equal_split_cost_terminates"),
    "split_le": ([("if (cost < best) { *out_axis = axis;", "if (cost <= best) { *out_axis = axis;")],
                 "test_plane_tie_goes_to_the_first_plane"),
    "axes_reversed": ([("for (uint32_t axis = 0; axis < 3; ++axis) {\n        float lo = cmin[axis]",
                        "for (uint32_t axis_r = 0; axis_r < 3; ++axis_r) {\n        uint32_t axis = 2 - axis_r;\n"
                        "        float lo = cmin[axis]")],
                      "test_axis_tie_goes_to_the_first_axis"),
    "planes_reversed": ([("for (uint32_t i = 0; i < B - 1; ++i) {\n            float cost = b->left_area[i] + b->right_area[i];",
                          "for (uint32_t ir = 0; ir < B - 1; ++ir) {\n            uint32_t i = B - 2 - ir;\n"
                          "            float cost = b->left_area[i] + b->right_area[i];")],
                        "test_plane_tie_goes_to_the_first_plane"),
    "factor_third": ([("* 0.3333f;", "* (1.0f / 3.0f);")], "test_centroid_factor_0_3333_decides_a_bin_boundary"),
    "stable_partition": ([("""        while (i <= j) {
            const float* c = &b->cen[3 * (size_t)b->idx[i]];
            int32_t bi = bin_of(c[axis], cmin[axis], scale, b->bins);
            if ((uint32_t)bi < split) {
                i++;
            } else {
                uint32_t t = b->idx[i]; b->idx[i] = b->idx[j]; b->idx[j] = t; j--;
            }
        }""", """        {
            uint32_t n0 = node->prim_count, lo0 = node->left_first, w = 0;
            uint32_t* tmp = (uint32_t*)malloc(n0 * sizeof(uint32_t));
            for (int pass = 0; pass < 2; ++pass)
                for (uint32_t q = lo0; q < lo0 + n0; ++q) {
                    const float* c = &b->cen[3 * (size_t)b->idx[q]];
                    int left = (uint32_t)bin_of(c[axis], cmin[axis], scale, b->bins) < split;
                    if (left == (pass == 0)) tmp[w++] = b->idx[q];
                    if (pass == 0 && left) i++;
                }
            memcpy(b->idx + lo0, tmp, n0 * sizeof(uint32_t));
            free(tmp);
            (void)j;
        }""")], "test_partition_is_the_swap_loop"),
    "right_first": ([("""        stack[sp++] = ri;   /* popped after the whole left subtree */
        stack[sp++] = li;""", """        stack[sp++] = li;
        stack[sp++] = ri;""")], "test_children_are_numbered_depth_first"),
    "breadth_first": ([("    stack[sp++] = root;", "    uint32_t head = 0;\n    stack[sp++] = root;"),
                       ("    while (sp) {\n        uint32_t ni = stack[--sp];",
                        "    while (head < sp) {\n        uint32_t ni = stack[head++];"),
                       ("""        stack[sp++] = ri;   /* popped after the whole left subtree */
        stack[sp++] = li;""", """        stack[sp++] = li;
        stack[sp++] = ri;""")], "test_children_are_numbered_depth_first"),
    "tlas_follow_moved": ([("if (metric == ORC_TLAS_UNION && A == count - 1) A = B;", "if (A == count - 1) A = B;")],
                          "test_tlas_merges_with_the_stale_index"),
    "tlas_ties_last": ([("if (area < smallest) { smallest = area; best = B; }",
                         "if (area <= smallest) { smallest = area; best = B; }")],
                       "test_tlas_match_ties_go_to_the_first_index"),
}

RUNNER = """
import sys
import numpy as np
sys.path.insert(0, {tests!r}); sys.path.insert(0, {root!r})
import oracle_api as O
O.ORACLE_PATH = {so!r}
import test_oracle_kat_builders as K
K.build = lambda tris, bins=8: O.blas_build(np.ascontiguousarray(np.array(tris, np.float32).reshape(-1, 9)), bins)
K.tlas = lambda inst: O.tlas_build(inst)
fails = []
for n in sorted(dir(K)):
    if n.startswith("test_") and "mutual" not in n:
        try:
            getattr(K, n)()
        except AssertionError:
            fails.append(n)
print(",".join(fails))
"""


def _build(tmp, flips):
    src = open(os.path.join(ORACLE, "bvh.c")).read()
    for a, b in flips:
        assert a in src, f"flip target not found in oracle/bvh.c: {a[:60]!r}"
        src = src.replace(a, b)
    for f in ("trace.c", "camera.c", "render.c", "adjoint.c", "post.c", "oracle.h"):
        with open(os.path.join(tmp, f), "w") as out:
            out.write(open(os.path.join(ORACLE, f)).read())
    with open(os.path.join(tmp, "bvh.c"), "w") as out:
        out.write(src)
    so = os.path.join(tmp, "liboracle.so")
    srcs = [os.path.join(tmp, f) for f in ("bvh.c", "trace.c", "camera.c", "render.c", "adjoint.c", "post.c")]
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-ffp-contract=off", "-shared", "-o", so] + srcs + ["-lm", "-lpthread"],
                   check=True, capture_output=True)
    return so


def _failing_kats(so):
    code = RUNNER.format(tests=os.path.join(ROOT, "tests"), root=ROOT, so=so)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return set(filter(None, r.stdout.strip().split(",")))


def test_unflipped_copy_passes_every_kat(tmp_path):
    assert _failing_kats(_build(str(tmp_path), [])) == set()


@pytest.mark.parametrize("flip", sorted(FLIPS))
def test_flipped_rule_fails_its_kat(tmp_path, flip):
    subs, kat = FLIPS[flip]
    fails = _failing_kats(_build(str(tmp_path), subs))
    assert kat in fails, (flip, fails)
