"""Write tests/golden/param_tables.json: fingerprints of the three literal lookup tables in the
reference's param.py (lines 4-6: arr_gparity, arr_encoder, arr_sgn).

The tables are parsed as text (the reference is not imported) and only their length and a SHA-256
of the canonical serialisation ",".join(str(v)) are stored, so that tests can check the closed forms
of gparity / encoder / sgn against all 65,536 entries without committing the tables themselves.
Run in the build container only:  python tests/golden/make_param_tables.py /root/reference/param.py
"""
import hashlib
import json
import os
import re
import sys


def fingerprint(values):
    s = ",".join(str(int(v)) for v in values)
    return hashlib.sha256(s.encode()).hexdigest()


def main(path):
    out = {}
    with open(path) as f:
        for line in f:
            m = re.match(r"\s*(arr_gparity|arr_encoder|arr_sgn)\s*=\s*np\.array\(\[(.*)\]\)", line)
            if m:
                vals = [int(x) for x in m.group(2).split(",") if x.strip()]
                out[m.group(1)] = {"length": len(vals), "sha256": fingerprint(vals)}
    if set(out) != {"arr_gparity", "arr_encoder", "arr_sgn"}:
        raise SystemExit("tables not found in %s" % path)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "param_tables.json")
    with open(dst, "w") as f:
        json.dump({"source": "param.py:4-6", "serialisation": "','.join(str(v))", "tables": out}, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/param.py")
