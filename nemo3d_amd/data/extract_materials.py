"""Extract the material records of NEMO-3D input decks into materials.json.

The reference keeps its material database inside every XML input deck
(``<group><cTag>Material</cTag> ... </group>``, read by
``find_Material_struct`` in ``nemo3d/src/io/i_mat.c:184-224``, first match
wins).  This script copies the numeric records (``mb_pv``, ``mb_strain``,
``Ev_offset``, ``unstrnd_cubic_cell_length`` and the Keating constants) of
the decks named on the command line into ``materials.json`` beside it, in
deck order, so that the library never needs the decks at run time.

Usage: python extract_materials.py DECK.xml [DECK2.xml ...]
"""
import json
import os
import sys
import xml.etree.ElementTree as ET

FIELDS_REAL = ["Ec", "Ev", "unstrnd_cubic_cell_length", "strain_alpha",
               "strain_beta", "Ev_offset", "tkel", "xval", "yval"]
FIELDS_ARRAY = ["mb_pv", "mb_strain", "strain_VFF"]
FIELDS_STR = ["material", "band_model", "comp1_mat", "comp2_mat"]


def _array(text):
    return [float(v) for v in text.replace("{", " ").replace("}", " ").split()]


def records(path):
    root = ET.parse(path).getroot()
    out = []
    for g in root.iter("group"):
        if (g.findtext("cTag") or "").strip() != "Material":
            continue
        rec = {"source": os.path.basename(path)}
        for p in g.findall("param"):
            tag = (p.findtext("cTag") or "").strip()
            val = (p.findtext("value") or "").strip()
            if tag in FIELDS_REAL:
                rec[tag] = float(val)
            elif tag in FIELDS_ARRAY:
                rec[tag] = _array(val)
            elif tag in FIELDS_STR:
                rec[tag] = val
        out.append(rec)
    return out


def main(argv):
    recs = []
    for path in argv:
        recs.extend(records(path))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "materials.json")
    with open(dst, "w") as f:
        json.dump({"records": recs}, f, indent=1)
    print("wrote %d records to %s" % (len(recs), dst))


if __name__ == "__main__":
    main(sys.argv[1:])
