"""Writes reference_tables.json.gz: the reference's generated data tables
(universe-config/gen/large-molecules.json, complexification.json,
complexes.json), re-serialised compactly as inputs of the reference-scene
tests, so that those tests run where the reference checkout is absent.  Data
only: particle types, interaction rows and complex layouts.

Run from the repository root: python3 tests/golden/make_reference_tables.py
"""
import gzip
import json
import os

GEN = "/root/reference/universe-config/gen"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json.gz")


def main():
    doc = {}
    for key, name in (("types", "large-molecules.json"), ("interactions", "complexification.json"),
                      ("complexes", "complexes.json")):
        with open(os.path.join(GEN, name)) as f:
            doc[key] = json.load(f)
    data = json.dumps(doc, separators=(",", ":"), sort_keys=True).encode()
    with gzip.GzipFile(OUT, "wb", mtime=0) as f:
        f.write(data)
    print(OUT, len(data), "bytes before compression")


if __name__ == "__main__":
    main()
