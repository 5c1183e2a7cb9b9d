"""Write tests/golden/rng_golden.json from the reference's own src/math/random.h.

The header is compiled where it lies in the reference tree into
oracle/_ref/libref_random.so (make -C oracle ref); this script calls it and
stores inputs and outputs only.  Run in a container that has the reference:

    make -C oracle ref && python tests/golden/make_rng_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle.oracle import ref_lib  # noqa: E402


def main() -> None:
    L = ref_lib()
    if L is None:
        raise SystemExit("oracle/_ref/libref_random.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(1234)
    seeds4 = [(0, 0, 0, 0), (1, 2, 3, 42), (-1, -1, -1, 42), (127, 63, 15, 42),
              (-2147483648, 2147483647, 0, 4294967295)]
    seeds4 += [tuple(int(x) for x in rng.integers(-5000, 5000, 3)) + (int(rng.integers(0, 2**32)),)
               for _ in range(40)]
    ms4 = [[list(s), int(L.ref_make_seed4(*s))] for s in seeds4]
    seeds5 = [(0, 0, 0, 42, 0), (12, 34, 5, 42, 1), (255, 255, 63, 42, 9)]
    seeds5 += [tuple(int(x) for x in rng.integers(0, 1024, 3)) + (42, int(rng.integers(0, 12)))
               for _ in range(40)]
    ms5 = [[list(s), int(L.ref_make_seed5(*s))] for s in seeds5]

    def draws(fn, s, n=12):
        out = np.zeros(n, np.uint32)
        fn(C.c_uint64(s), n, out.ctypes.data_as(C.c_void_p))
        return [int(x) for x in out]

    ctor_seeds = [0, 1, 42, 2**32 - 1, 2**63 + 12345] + [int(x) for x in rng.integers(0, 2**62, 20)]
    ctor = [[s, draws(L.ref_random_ctor_draws, s)] for s in ctor_seeds]
    state = [[s, draws(L.ref_random_state_draws, s)] for s in ctor_seeds]
    uints = [0, 1, 511, 512, 2**31 - 1, 2**31, 2**32 - 1] + [int(x) for x in rng.integers(0, 2**32, 30)]
    norm = [[u, float(np.float32(L.ref_normalized_uint(u)))] for u in uints]
    bern = [[float(x), float(L.ref_bernoulli(float(x)))] for x in
            [0.0, 0.25, 0.4999999, 0.5, 0.5000001, 0.75, 0.9999999]]
    out = {
        "source": "reference src/math/random.h compiled into oracle/_ref/libref_random.so",
        "make_seed4": ms4, "make_seed5": ms5, "random_ctor_draws": ctor,
        "random_setstate_draws": state, "normalized_uint": norm, "bernoulli": bern,
    }
    path = Path(__file__).with_name("rng_golden.json")
    path.write_text(json.dumps(out, indent=1))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
