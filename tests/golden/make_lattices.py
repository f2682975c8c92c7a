"""Write tests/golden/lattices.json from the reference's built-in lattice YAML files.

Run in a container where the reference tree is mounted at /root/reference:
    python tests/golden/make_lattices.py
The YAML files are data (src/isle/resources/lattices/*.yml, MIT licence); this script
reads them with yaml.SafeLoader (the custom !lattice tag is mapped to a plain dict) and
stores, per lattice, nx, the adjacency list and the hopping strengths.
"""
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
        adj = [list(map(int, a)) for a in d["adjacency"]]
        hop = d["hopping"]
        if not isinstance(hop, list):
            hop = [hop] * len(adj)
        out[path.stem] = {"nx": len(d["positions"]), "adjacency": adj,
                          "hopping": [float(h) for h in hop], "name": d.get("name", "")}
    OUT.write_text(json.dumps(out, indent=None, separators=(",", ":")))
    print("wrote", OUT, len(out), "lattices")


if __name__ == "__main__":
    main()
