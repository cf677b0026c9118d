"""Writes reference_tables.json: the numeric tables the reference's forward model reads
(model/samodel.c:103-286), parsed from the reference's source text.

Run in a tree that holds the reference (it is read as text, nothing is compiled or run):
    python3 tests/golden/make_reference_tables.py /root/reference/model/samodel.c
The output is data only (name -> list of numbers and the source line it was read from)."""
import json
import os
import re
import sys

NAMES = {
    # name in this fixture: (array name in model/samodel.c)
    "ref_wl": "ref_wlens",
    "aw": "aw_Pope_Fry1997",
    "bbw": "bbw_Morel_1974",
    "a0": "a0_Lee",
    "a1": "a1_Lee",
    # the bottom table in the order model/samodel.c:466-478 fills it
    "bottom_0_sand": "bottom_type_sand",
    "bottom_1_seagrass": "bottom_type_seagrass",
    "bottom_2_coral": "bottom_type_coral",
    "bottom_3_macrophytes": "bottom_type_macrophytes",
    "bottom_4_dark_sediment": "bottom_type_dark_sediment",
    "bottom_5_coral_sand": "bottom_type_coral_sand",
    "bottom_6_green_algae": "bottom_type_green_algae",
    "bottom_7_red_algae": "bottom_type_red_algae",
}


def main(path):
    text = open(path).read()
    out = {}
    for key, name in NAMES.items():
        m = re.search(r"double\s+%s\s*\[\s*\]\s*=\s*\{([^}]*)\}" % re.escape(name), text)
        if m is None:
            raise SystemExit("table %s not found" % name)
        line = text[:m.start()].count("\n") + 1
        body = re.sub(r"//[^\n]*", "", m.group(1))
        vals = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
        if len(vals) != 41:
            raise SystemExit("table %s has %d entries" % (name, len(vals)))
        out[key] = {"source": "model/samodel.c:%d" % line, "values": vals}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/model/samodel.c")
