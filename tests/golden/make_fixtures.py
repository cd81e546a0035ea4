"""Writes tests/golden/fixtures.json from the reference's validation data.

The reference keeps 28 small networks (validation/fixtures/<name>/edges.csv
and meta.json) and the values its oracles produced for them
(validation/golden/{definition,networkx}/<name>/<variant>.csv, one
``%.17g`` value per line under a ``value`` header; see
validation/common.py:341-352; the centrality_extra variants write NaN as
-999, the topography ones as -1). This script copies the edge lists, the
metadata (with the node types of two-mode fixtures) and the goldens of
betweenness, edge betweenness, betweenness for sources and targets,
closeness / harmonic centrality, Freeman centralization and centrality
normalized for two-mode data into one JSON file, as data. Nothing of the reference is imported or run.

Run once in a tree where the reference is checked out:

    python tests/golden/make_fixtures.py /root/reference

The output is committed; tests read only the JSON file.
"""
import csv
import json
import sys
from pathlib import Path

VARIANTS = [
    "betweenness",
    "betweenness_normalized",
    "betweenness_weighted",
    "betweenness_weighted_strength",
    "betweenness_networkx",
    "betweenness_networkx_weighted",
    "edge_betweenness",
    "edge_betweenness_normalized",
    "edge_betweenness_weighted",
    "edge_betweenness_weighted_strength",
    # betweenness for sources and targets (validation/manifest.d/centrality_extra.yaml)
    "cx_bc_all_distinct",
    "cx_bc_all_distinct_weighted",
    "cx_ebc_all_distinct",
    "cx_bc_subset_nx",
    "cx_bc_subset_nx_normalized",
    "cx_bc_subset_nx_weighted",
    "cx_ebc_subset_nx",
    "cx_ebc_subset_nx_weighted",
    "cx_ebc_subset_proportional",
    "cx_bc_sampled",
    "cx_bc_sampled_raw",
    "cx_bc_sampled_endpoints",
    "cx_bc_sampled_weighted",
    "cx_ebc_sampled",
    # closeness and harmonic centrality (validation/manifest.d/topography.yaml; NaN as -1)
    "closeness",
    "closeness_normalized",
    "closeness_in",
    "closeness_all",
    "closeness_wf",
    "closeness_sna",
    "closeness_ucinet",
    "closeness_ucinet_in",
    "closeness_weighted",
    "closeness_wf_weighted",
    "closeness_weighted_strength",
    "harmonic",
    "harmonic_normalized",
    "harmonic_in",
    "harmonic_weighted",
    # Freeman centralization (topography.yaml; one value, NaN as -1)
    "centralization_degree",
    "centralization_degree_in",
    "centralization_degree_sna",
    "centralization_degree_ucinet",
    "centralization_betweenness",
    "centralization_closeness",
    "centralization_closeness_sna",
    "centralization_closeness_ucinet",
    # centrality normalized for two-mode data (validation/manifest.d/twomode.yaml)
    "twomode_degree",
    "twomode_degree_raw",
    "twomode_closeness",
    "twomode_closeness_ignore",
    "twomode_closeness_raw",
    "twomode_betweenness",
    "twomode_betweenness_raw",
]
ORACLES = ["definition", "networkx"]


def read_values(path: Path):
    with open(path, newline="") as fh:
        rows = list(csv.reader(fh))
    assert rows[0] == ["value"], path
    out = []
    for r in rows[1:]:
        v = r[0]
        out.append(None if v in ("NA", "") else float(v))
    return out


def main(ref: Path, dest: Path) -> None:
    val = ref / "validation"
    fixtures = {}
    for d in sorted((val / "fixtures").iterdir()):
        if not (d / "meta.json").exists():
            continue
        meta = json.loads((d / "meta.json").read_text())
        src, tgt, w = [], [], []
        with open(d / "edges.csv", newline="") as fh:
            for row in csv.DictReader(fh):
                src.append(int(row["from"]))
                tgt.append(int(row["to"]))
                if meta["weighted"]:
                    w.append(float(row["weight"]))
        golden = {}
        for oracle in ORACLES:
            for var in VARIANTS:
                p = val / "golden" / oracle / d.name / f"{var}.csv"
                if p.exists():
                    golden.setdefault(var, {})[oracle] = read_values(p)
        fixtures[d.name] = {
            "n": meta["n"],
            "directed": meta["directed"],
            "weighted": meta["weighted"],
            "tags": meta["tags"],
            "src": src,
            "tgt": tgt,
            "weights": w if meta["weighted"] else None,
            "types": meta.get("type"),
            "golden": golden,
        }
    version = (val / "golden" / "networkx" / "VERSION").read_text().strip()
    doc = {
        "about": "knappnet validation fixtures and betweenness goldens, copied as data",
        "networkx_version": version,
        "tolerance": {"rel": 1e-9, "abs": 1e-9, "abs_normalized": 1e-12},
        "fixtures": fixtures,
    }
    dest.write_text(json.dumps(doc, separators=(",", ":")) + "\n")
    print(f"wrote {len(fixtures)} fixtures to {dest}")


if __name__ == "__main__":
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    main(ref, Path(__file__).resolve().parent / "fixtures.json")
