"""Writes tests/golden/<deck>_deck.json: the parsed input decks of the
reference's own test runs (NEMO_3D/nemo3d/example/TestRun/inputs/
InGaAs_QD.xml, whose eigenvalues TestRun/InGaAs_QD.old_nd_Ek_ascii holds,
and PostNEMO3D/bin/test/silicon4.xml, whose atoms silicon4.nd_rAtom /
nd_nbrIndx hold), read by nemo3d_amd.deck.  The material list keeps only
the records the deck's shapes use (the first record of each compound and
band model, as find_Material_struct picks them), so the fixtures stay
small.  Run in a container that has the reference tree:
  python tests/golden/make_deck_fixture.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from nemo3d_amd import deck, materials  # noqa: E402

REF = "/root/reference/NEMO_3D"
DECKS = {
    "ingaas_qd": REF + "/nemo3d/example/TestRun/inputs/InGaAs_QD.xml",
    "silicon4": REF + "/PostNEMO3D/bin/test/silicon4.xml",
}


def prune(tree):
    band = deck.find(tree, "band_model")
    used = []
    for comp in deck.compounds(tree):
        rec = materials.find_material(comp, band, records=deck.material_records(tree),
                                      spinless_view=True)
        used.append(rec)
    tree = json.loads(json.dumps(tree))
    tree["material_list"] = used
    return tree


def main():
    for name, path in DECKS.items():
        tree = prune(deck.read(path))
        dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), name + "_deck.json")
        with open(dst, "w") as f:
            json.dump(tree, f, indent=0, sort_keys=False)
        print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main()
