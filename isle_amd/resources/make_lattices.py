"""Write isle_amd/resources/lattices.json, the data behind isle_amd.LATTICES, from the
reference's built-in lattice files (src/isle/resources/lattices/*.yml, MIT licence; data,
not code).  Run where the reference tree is mounted at /root/reference:
    python isle_amd/resources/make_lattices.py
Each file is read with yaml.SafeLoader (the !lattice tag mapped to a plain mapping) and
stored as {name, comment, adjacency, hopping, positions}, hopping expanded to one strength
per bond."""
import json
import pathlib

import yaml

SRC = pathlib.Path("/root/reference/src/isle/resources/lattices")
OUT = pathlib.Path(__file__).resolve().parent / "lattices.json"


class _Loader(yaml.SafeLoader):
    pass


_Loader.add_constructor("!lattice", lambda loader, node: loader.construct_mapping(node, deep=True))


def main():
    out = {}
    for path in sorted(SRC.glob("*.yml")):
        d = yaml.load(path.read_text(), Loader=_Loader)
        adjacency = [[int(i), int(j)] for i, j in d["adjacency"]]
        hopping = d["hopping"]
        if not isinstance(hopping, list):
            hopping = [hopping] * len(adjacency)
        out[path.stem] = {"name": d.get("name", ""), "comment": d.get("comment", ""), "adjacency": adjacency,
                          "hopping": [float(h) for h in hopping],
                          "positions": [[float(c) for c in p] for p in d["positions"]]}
    OUT.write_text(json.dumps(out, separators=(",", ":")))
    print("wrote", OUT, len(out), "lattices")


if __name__ == "__main__":
    main()
