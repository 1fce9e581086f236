"""Write tests/golden/e2e_sanity.npz from the reference's end-to-end golden.

The reference test tests/test_pgmax.py:35-421 hard-codes the 84 ftov messages
PGMax produces after 100 max-product iterations (lines 63-152) and the 12 MAP
states (lines 252-265). This script reads those two literals out of the
reference source with ``ast`` (nothing is imported or executed) and stores
them as data. The graph and its evidence are rebuilt by the test itself.

Usage: python tests/golden/make_e2e_sanity.py [/root/reference]
"""

import ast
import os
import sys

import numpy as np


def main():
  ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
  src = open(os.path.join(ref, "tests", "test_pgmax.py")).read()
  tree = ast.parse(src)
  msgs = None
  map_states = None
  for node in ast.walk(tree):
    if isinstance(node, ast.Assign) and len(node.targets) == 1:
      name = getattr(node.targets[0], "id", None)
      if name == "true_final_msgs_output":
        lst = node.value.args[0].args[0]  # jax.device_put(jnp.array([...]))
        msgs = np.array(ast.literal_eval(lst), dtype=np.float32)
      if name == "true_map_state_output":
        entries = []
        for k, v in zip(node.value.keys, node.value.values):
          group = k.elts[0].id
          index = ast.literal_eval(k.elts[1])
          entries.append((group, index, ast.literal_eval(v)))
        map_states = entries
  assert msgs is not None and msgs.shape == (84,)
  assert map_states is not None and len(map_states) == 12
  out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "e2e_sanity.npz")
  np.savez(
      out,
      final_msgs=msgs,
      map_group=np.array([g for g, _, _ in map_states]),
      map_index=np.array([i for _, i, _ in map_states], dtype=np.int64),
      map_state=np.array([s for _, _, s in map_states], dtype=np.int64),
  )
  print("wrote", out)


if __name__ == "__main__":
  main()
