"""Write tests/golden/ref_tables.npz: sampled entries of the reference's own
data tables (data only; no reference code is imported or run).

  SHT_ring_skip_table/NNNN.bin   int32 ring_tab(m), m = 0..len-1
  plm_table/plm_pos_NNNN.bin     int32 pos(i, m), (2*nside, 4*nside+1) Fortran order
  plm_table/plm_val{1,2}_NNNN.bin float64 P_{pos,m}(theta_i), P_{pos+1,m}(theta_i)

The tables are sampled (seeded) rather than copied whole.  Run in the
container that holds the reference:  python tests/golden/make_golden.py
"""
import os

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_tables.npz")
SEED = 20261203


def main():
    rng = np.random.default_rng(SEED)
    out = {"seed": np.array(SEED)}
    for ns in [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048]:
        rt = np.fromfile(f"{REF}/SHT_ring_skip_table/{ns:04d}.bin", dtype=np.int32)
        k = min(len(rt), 64)
        m = np.unique(np.concatenate([rng.choice(len(rt), k, replace=False), [0, len(rt) - 1]]))
        out[f"rt_{ns}_len"] = np.array(len(rt))
        out[f"rt_{ns}_m"] = m.astype(np.int32)
        out[f"rt_{ns}_v"] = rt[m]
    for ns in [1, 2, 4, 8, 16, 32, 64, 128, 256]:
        pos = np.fromfile(f"{REF}/plm_table/plm_pos_{ns:04d}.bin", dtype=np.int32)
        nm = len(pos) // (2 * ns)
        pos = pos.reshape(2 * ns, nm, order="F")
        k = min(pos.size, 400)
        flat = rng.choice(pos.size, k, replace=False)
        i, m = np.unravel_index(flat, pos.shape)
        out[f"plm_{ns}_i"] = i.astype(np.int32)
        out[f"plm_{ns}_m"] = m.astype(np.int32)
        out[f"plm_{ns}_pos"] = pos[i, m]
        if ns <= 128:
            v1 = np.fromfile(f"{REF}/plm_table/plm_val1_{ns:04d}.bin").reshape(2 * ns, nm, order="F")
            v2 = np.fromfile(f"{REF}/plm_table/plm_val2_{ns:04d}.bin").reshape(2 * ns, nm, order="F")
            out[f"plm_{ns}_v1"] = v1[i, m]
            out[f"plm_{ns}_v2"] = v2[i, m]
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
