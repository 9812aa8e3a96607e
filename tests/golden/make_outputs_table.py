"""Writes outputs_table.json: the reference's output-file definitions as data.

Reads MainSrc/OUTPUTS.f08 of the reference (as text) and records, for OUTFILE_INIT,
OUTFILE_INIT_Tgen and OUTFILE_VARDEFS, every dimension (NF_DEF_DIM), global attribute
(nf90_put_att on NF90_GLOBAL), variable (NF_DEF_VAR: name, type, dimension names in
Fortran order), its Units (NF_DEF_UNIT) and other attributes, each with the IF
conditions it is defined under.  The tests compare the library's files against this
table.  Usage: python tests/golden/make_outputs_table.py [/path/to/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "MainSrc", "OUTPUTS.f08")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "outputs_table.json")


def statements(body):
    """Fortran statements with continuations joined and comments removed."""
    out, cur = [], ""
    for raw in body.splitlines():
        line, q = [], None
        for ch in raw:
            if q:
                line.append(ch)
                if ch == q:
                    q = None
            elif ch in "'\"":
                q = ch
                line.append(ch)
            elif ch == "!":
                break
            else:
                line.append(ch)
        s = "".join(line).strip()
        if not s:
            continue
        if s.endswith("&"):
            cur += s[:-1].strip()
            continue
        out.append(cur + s)
        cur = ""
    return out


def split_args(s):
    args, depth, cur, q = [], 0, "", None
    for ch in s:
        if q:
            cur += ch
            if ch == q:
                q = None
            continue
        if ch in "'\"":
            q = ch
        elif ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            args.append(cur.strip())
            cur = ""
        else:
            cur += ch
    args.append(cur.strip())
    return args


def literal(a):
    a = a.strip()
    if a.startswith("'") and a.endswith("'"):
        return a[1:-1]
    if re.fullmatch(r"-?\d+", a):
        return int(a)
    return {"expr": a}


def parse(sub, src):
    m = re.search(r"SUBROUTINE\s+%s\s*\((.*?)END SUBROUTINE %s" % (sub, sub), src, re.S)
    body = m.group(0)
    conds, dims, gatts, vars_, ids = [], [], [], [], {}
    for st in statements(body):
        m = re.match(r"IF\s*\((.*)\)\s*THEN$", st, re.I)
        if m:
            conds.append(m.group(1).strip())
            continue
        m = re.match(r"ELSE\s*IF\s*\((.*)\)\s*THEN$", st, re.I)
        if m:
            conds[-1] = m.group(1).strip()
            continue
        if re.match(r"ELSE$", st, re.I):
            conds[-1] = "NOT " + conds[-1]
            continue
        if re.match(r"END\s*IF$", st, re.I):
            conds.pop()
            continue
        here = list(conds)
        m = re.match(r"IF\s*\((.*?)\)\s*(CALL .*|Status.*)$", st, re.I)
        if m:
            here.append(m.group(1).strip())
            st = m.group(2)
        m = re.match(r"CALL\s+NF_DEF_DIM\((.*)\)$", st)
        if m:
            a = split_args(m.group(1))
            dims.append({"id": a[0], "name": literal(a[2]), "len": a[3], "cond": here})
            continue
        m = re.match(r"CALL\s+NF_DEF_VAR\((.*)\)$", st)
        if m:
            a = split_args(m.group(1))
            d = a[2]
            dl = [] if d == "Z" else [x.strip() for x in d.strip("(/)").split(",")]
            v = {"id": a[0], "name": literal(a[3]), "type": literal(a[4]), "dims": dl, "cond": here, "atts": {}}
            vars_.append(v)
            ids[a[0]] = v
            continue
        m = re.match(r"CALL\s+NF_DEF_UNIT\((.*)\)$", st)
        if m:
            a = split_args(m.group(1))
            ids[a[1]]["atts"]["Units"] = literal(a[2])
            continue
        m = re.match(r"Status\s*=\s*nf90_put_att\((.*)\)$", st)
        if m:
            a = split_args(m.group(1))
            if a[1] == "NF90_GLOBAL":
                gatts.append({"name": literal(a[2]), "value": literal(a[3]), "cond": here})
            else:
                ids[a[1]]["atts"][literal(a[2]) if not isinstance(literal(a[2]), dict) else a[2]] = literal(a[3])
    for v in vars_:
        del v["id"]
    return {"dims": dims, "gatts": gatts, "vars": vars_}


def main():
    src = open(SRC).read()
    table = {s: parse(s, src) for s in ("OUTFILE_INIT", "OUTFILE_INIT_Tgen", "OUTFILE_VARDEFS")}
    # dimension-id variables -> dimension names, from OUTFILE_INIT
    table["dim_ids"] = {d["id"]: d["name"] for d in table["OUTFILE_INIT"]["dims"]}
    with open(OUT, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    n = sum(len(table[s]["vars"]) for s in ("OUTFILE_INIT", "OUTFILE_INIT_Tgen", "OUTFILE_VARDEFS"))
    print(f"wrote {OUT}: {n} variable definitions")


if __name__ == "__main__":
    main()
