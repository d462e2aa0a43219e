"""Writes the fixtures of the KTT-JSON -> T4 results conversion from the data files the reference
holds for it:

* ``t4_vector_add.json``: Tutorials/03KernelTuning/SimulationInput.json (4 results of a vectorAdd
  search on an RTX 2080 Ti, with profiling counters) and its conversion
  Tutorials/03KernelTuning/SimulationInput_converted.json, as {"ktt": ..., "t4": ...}.
* ``t4_coulomb2080_head.json``: the first 12 results of
  Examples/CoulombSum3d/coulomb_2080_full_search_space.json (33 counters each) and the same 12
  results of coulomb_2080_full_search_space_converted.json.

Both pairs were made by the reference's Scripts/ConvertJsonOutput.py; the files are copied as
data, unchanged apart from the cut to 12 results.

Usage: python tests/golden/make_t4_fixtures.py [reference root]
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
HEAD = 12


def main() -> int:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    tutorial = root / "Tutorials/03KernelTuning"
    pair = {"ktt": json.loads((tutorial / "SimulationInput.json").read_text()),
            "t4": json.loads((tutorial / "SimulationInput_converted.json").read_text())}
    (HERE / "t4_vector_add.json").write_text(json.dumps(pair, indent=1) + "\n")

    example = root / "Examples/CoulombSum3d"
    ktt = json.loads((example / "coulomb_2080_full_search_space.json").read_text())
    t4 = json.loads((example / "coulomb_2080_full_search_space_converted.json").read_text())
    ktt["Results"] = ktt["Results"][:HEAD]
    t4["results"] = t4["results"][:HEAD]
    (HERE / "t4_coulomb2080_head.json").write_text(json.dumps({"ktt": ktt, "t4": t4}, indent=1) + "\n")
    print("wrote t4_vector_add.json and t4_coulomb2080_head.json")
    return 0


if __name__ == "__main__":
    sys.exit(main())
