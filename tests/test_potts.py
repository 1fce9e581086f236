"""GPU parity for batched uneven Potts grids (the 2,048-graph benchmark workload).

The per-phase shape runs square pairwise tables of 3..8 states on S lanes
per factor (k_pair_lanes); its messages and deltas must equal the FP32
oracle bit for bit: random tables with clipped (+-inf, 1e7) and NaN
log-potentials, random evidence and initial messages, every tile size. The
full 2,048-graph union is checked per graph on a sample of its graphs.
"""

import numpy as np
import pytest

from oracle import lbp
from pgmax_amd import _lib
from pgmax_amd import batch
from pgmax_amd import workloads

pytestmark = pytest.mark.gpu


def _np(x):
  return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


@pytest.fixture(autouse=True)
def _options(gpu):
  keys = ("mode", "fuse", "tile", "vec", "lanes", "xcd", "share")
  prev = {k: _lib.set_option(k, 1 if k in ("lanes", "share") else 0) for k in keys}
  yield
  for k, v in prev.items():
    _lib.set_option(k, v)


def _union_inputs(pb, rng, nonfinite):
  lp = rng.normal(0, 1, size=pb.log_potentials.shape).astype(np.float32)
  if nonfinite:
    idx = rng.choice(lp.size, size=12, replace=False)
    lp[idx[:3]] = np.inf
    lp[idx[3:6]] = -np.inf
    lp[idx[6:9]] = 1e7
    lp[idx[9:11]] = -3e8
    lp[idx[11]] = np.nan
  ev = rng.normal(0, 1, size=pb.evidence.shape).astype(np.float32)
  return lp, ev


@pytest.mark.parametrize("fuse,lanes,tile,xcd", [
    (0, 1, 64, 0), (0, 1, 128, 0), (0, 1, 256, 0), (0, 0, 256, 0), (0, 1, 256, 4), (0, 0, 128, 2),
    (1, 1, 256, 0), (1, 1, 64, 2), (1, 0, 128, 0)])
@pytest.mark.parametrize("temperature", [0.0, 1.0, 0.35])
@pytest.mark.parametrize("nonfinite", [False, True])
def test_pair_lanes_match_oracle(fuse, lanes, tile, xcd, temperature, nonfinite):
  """Nine uneven grids holding every state count 2..8, per-phase shape: bitwise
  equal to the oracle's messages and deltas over 9 iterations, on S lanes
  per factor with tiles of 1 / 2 / 4 passes per wavefront (several tiles
  per block), one factor per thread, with workgroups renumbered into runs of
  2 or 4 per XCD, and with the variable sums formed on the factor side (one
  launch per iteration: 2 x 2 and square blocks on lane groups, or one
  factor per thread)."""
  pb = workloads.potts_batch(num_graphs=9, seed=17, side_range=(9, 31), state_range=(2, 8))
  rng = np.random.default_rng(int(tile + 1000 * temperature) + nonfinite)
  lp, ev = _union_inputs(pb, rng, nonfinite)
  bb = batch.BatchedBP(pb.wirings, temperature=temperature)
  ft0 = rng.normal(0, 0.4, size=bb.union.num_edge_states).astype(np.float32)
  _lib.set_option("mode", 2)
  _lib.set_option("fuse", fuse)
  _lib.set_option("tile", tile)
  _lib.set_option("lanes", lanes)
  _lib.set_option("xcd", xcd)
  out, d = bb.run(lp, ft0, ev, 9, 0.5, temperature)
  ref, ref_d = lbp.run_with_diffs(lbp.graph_from_flat(bb.union), lp, ft0, ev, 9, 0.5,
                                  temperature)
  np.testing.assert_array_equal(_np(out), ref)
  np.testing.assert_array_equal(_np(d), ref_d)
  bb.close()


@pytest.mark.parametrize("fuse,tile", [(0, 64), (0, 256), (1, 128), (1, 256)])
@pytest.mark.parametrize("temperature", [0.0, 1.0, 0.35])
def test_shared_tables_match_oracle(fuse, tile, temperature):
  """Grids whose factors share one table (beta * I_S, as PairwiseFactorGroup
  builds from a 2-D matrix) are read as one table per lane tile (option
  "share"): bitwise equal to reading every factor's own table and to the
  oracle. Some grids' shared tables hold NaN, +inf or -1e7 (still shared);
  in one grid a single factor's table differs in one entry (its tile must
  not share), in another only by the sign of a zero (different bits). With
  fuse 1 the lane kernel forms the variable sums itself (one launch per
  iteration), 2 x 2 grids included."""
  pb = workloads.potts_batch(num_graphs=10, seed=23, side_range=(9, 31),
                             state_range=(2 if fuse else 3, 8))
  lay_lp = np.cumsum([0] + [w.num_log_potentials for w in pb.wirings])
  lp = pb.log_potentials.copy()
  s = pb.states
  def table_entries(g, r, c):  # entry (r, c) of every factor's table in graph g
    return np.arange(lay_lp[g] + r * s[g] + c, lay_lp[g + 1], s[g] * s[g])
  g3 = [g for g in range(pb.num_graphs) if s[g] >= 3]  # grids with a row and column 2
  assert len(g3) >= 5
  lp[table_entries(g3[0], 0, 1)] = np.nan
  lp[table_entries(g3[1], 1, 1)] = np.inf
  lp[table_entries(g3[2], 2, 0)] = -1e7
  lp[table_entries(g3[3], 0, 2)[7]] = 0.25  # one factor differs
  lp[table_entries(g3[4], 1, 2)] = 0.0
  lp[table_entries(g3[4], 1, 2)[3]] = -0.0  # one factor differs in the sign of a zero
  rng = np.random.default_rng(31)
  ev = rng.normal(0, 1, size=pb.evidence.shape).astype(np.float32)
  bb = batch.BatchedBP(pb.wirings, temperature=temperature)
  ft0 = rng.normal(0, 0.4, size=bb.union.num_edge_states).astype(np.float32)
  _lib.set_option("mode", 2)
  _lib.set_option("fuse", fuse)
  _lib.set_option("tile", tile)
  outs = []
  prev = _lib.set_option("share", 1)
  try:
    for share in (1, 0):
      _lib.set_option("share", share)
      out, d = bb.run(lp, ft0, ev, 10, 0.5, temperature)
      outs.append((_np(out), _np(d)))
  finally:
    _lib.set_option("share", prev)
  np.testing.assert_array_equal(outs[0][0], outs[1][0])
  np.testing.assert_array_equal(outs[0][1], outs[1][1])
  ref, ref_d = lbp.run_with_diffs(lbp.graph_from_flat(bb.union), lp, ft0, ev, 10, 0.5,
                                  temperature)
  np.testing.assert_array_equal(outs[0][0], ref)
  np.testing.assert_array_equal(outs[0][1], ref_d)
  bb.close()


def test_potts_2048_full_size_sampled_graphs():
  """The full benchmark union (2,048 graphs, 2.58e8 edge states): 3 iterations
  in the tuned shape; the messages of sampled graphs (the largest of each
  state count, plus the first and last graph) equal the oracle run on that
  graph alone, bit for bit (graphs are disjoint, so each graph's messages
  depend on its own inputs only)."""
  import torch  # pylint: disable=g-import-not-at-top
  pb = workloads.potts_batch(2048, seed=0)
  bb = batch.BatchedBP(pb.wirings, temperature=1.0)
  lp = torch.as_tensor(pb.log_potentials, device="cuda")
  ev = torch.as_tensor(pb.evidence, device="cuda")
  bb.autotune(lp, None, ev, 0.5, 1.0, probe_iters=2)
  out, _ = bb.run(lp, None, ev, 3, 0.5, 1.0)
  out = _np(out)
  size = pb.sides.astype(np.int64) ** 2 * pb.states
  sample = {0, pb.num_graphs - 1}
  for s in range(2, 9):
    sel = np.flatnonzero(pb.states == s)
    if sel.size:
      sample.add(int(sel[np.argmax(size[sel])]))
  lay = bb.layout
  for g in sorted(sample):
    w = pb.wirings[g]
    lp_g = pb.log_potentials[lay.lp_index[g][0]:lay.lp_index[g][-1] + 1]
    ev_g = pb.evidence[lay.vs_offsets[g]:lay.vs_offsets[g + 1]]
    ref, _ = lbp.run_with_diffs(lbp.graph_from_flat(w), lp_g,
                                np.zeros(w.num_edge_states, np.float32), ev_g, 3, 0.5, 1.0)
    np.testing.assert_array_equal(out[lay.es_index[g]], ref,
                                  err_msg=f"graph {g}: side {pb.sides[g]}, {pb.states[g]} states")
  bb.close()
