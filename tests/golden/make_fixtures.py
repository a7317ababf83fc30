"""Extracts golden data from the reference's published result files.

Reads text tables under /root/reference/Results (never code) and writes small
JSON fixtures next to this script:

* bibd_counts.json   per 2-(v,k,lambda): number of nonisomorphic designs, the
                     |Aut| histogram and the "fully constructed" count, from
                     Results/BIBDs/V = <v>/<v>_<k>_<lambda>.txt and
                     Results/BIBDs/Partial Results/25_4_1_tmp.txt.
* printed_bibds.json canonical BIBD matrices printed with |Aut(M)| in
                     Results/Quasi-Combined_BIBDs/V = <v>/*.txt.
* format_golden.json verbatim |Aut(D)| tables and printed matrix blocks
                     (expected output of the text formatters).
* kfact_k10.json     the 396 canonical 1-factorizations of K_10 with |Aut(M)|,
                     k-Factorization/Results/Complete_graphs/10/10x9x2/K_OrderedMatrices.txt,
                     the uniform K_12 factorization of
                     .../12/12x11x2_66/U_OrderedMatrices.txt, and the perfect
                     1-factorizations of K_10 and K_14 with |Aut(M)|,
                     .../10/10x9x2/P_OrderedMatrices.txt, .../14/14x13x2/P_OrderedMatrices.txt,
                     and the uniform ones of K_8 (cycles 4:4) and K_10 (4:6),
                     .../8/8x7x2_44/U_OrderedMatrices.txt, .../10/10x9x2_46/U_OrderedMatrices.txt;
                     k15x3 / p15x3: the 15-player k-systems in groups of three
                     (7 days of five triples) of .../15/15x7x3/K_OrderedMatrices.txt and
                     P_OrderedMatrices.txt (rows: the day's triples in order).

Run: python tests/golden/make_fixtures.py  (needs the reference checkout).
"""
import glob
import json
import os
import re

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def read(path):
    raw = open(path, "rb").read()
    if raw[:2] in (b"\xff\xfe", b"\xfe\xff"):
        return raw.decode("utf-16")
    return raw.decode("utf-8", errors="replace")


def parse_summary(text):
    """|Aut(D)| histogram table + counts of a BIBD result file."""
    m = re.search(r"BIBD\(\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\)", text)
    if not m:
        return None
    v, b, r, k, lam = map(int, m.groups())
    hist = {}
    in_table = False
    total = None
    for line in text.splitlines():
        if "|Aut(D)|" in line:
            in_table = True
            continue
        if not in_table:
            continue
        t = line.strip()
        if t.startswith("Total:"):
            total = int(t.split()[1])
            break
        nums = t.split()
        if len(nums) >= 2 and all(x.isdigit() for x in nums[:2]):
            hist[int(nums[0])] = hist.get(int(nums[0]), 0) + int(nums[1])
    fc = re.search(r"(\d+)\s+matri(?:x|ces)\s+(?:was|were)\s+fully constructed", text)
    return dict(v=v, b=b, r=r, k=k, **{"lambda": lam}, designs=total, aut_hist=hist,
                fully_constructed=int(fc.group(1)) if fc else None)


def bibd_counts():
    out = []
    for path in sorted(glob.glob(os.path.join(REF, "Results/BIBDs/V = *", "*.txt"))):
        s = parse_summary(read(path))
        if s and s["designs"] is not None:
            s["source"] = os.path.relpath(path, REF)
            out.append(s)
    tmp = os.path.join(REF, "Results/BIBDs/Partial Results/25_4_1_tmp.txt")
    if os.path.exists(tmp):
        s = parse_summary(read(tmp))
        if s:
            s["source"] = os.path.relpath(tmp, REF)
            out.append(s)
    return out


def parse_matrices(text):
    mats, cur, aut = [], None, None
    for line in text.splitlines():
        m = re.match(r"\s*Matrix #\s*\S+\s+\|Aut\(M\)\|\s*=\s*(\d+)", line)
        if m:
            if cur:
                mats.append((cur, aut))
            cur, aut = [], int(m.group(1))
            continue
        t = line.strip()
        if cur is not None and t and set(t) <= set("01"):
            cur.append(t)
        elif cur is not None and cur:
            mats.append((cur, aut))
            cur = None
    if cur:
        mats.append((cur, aut))
    return mats


def printed_bibds():
    out = []
    for path in sorted(glob.glob(os.path.join(REF, "Results/Quasi-Combined_BIBDs/V = *", "*.txt"))):
        text = read(path)
        head = re.search(r"BIBD\(\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\)", text)
        mats = parse_matrices(text)
        if not head or not mats:
            continue
        v, b, r, k, lam = map(int, head.groups())
        out.append(dict(v=v, b=b, r=r, k=k, **{"lambda": lam}, source=os.path.relpath(path, REF),
                        matrices=[dict(rows=rows, aut=a) for rows, a in mats]))
    return out


def parse_kfact(text, n):
    """Blocks of '<i>: |Aut(M)| = a' followed by rows of n player numbers."""
    mats, cur, aut = [], None, None
    for line in text.splitlines():
        m = re.match(r"\s*(\d+):\s*\|Aut\(M\)\|\s*=\s*(\d+)", line)
        if m:
            if cur:
                mats.append((cur, aut))
            cur, aut = [], int(m.group(2))
            continue
        t = line.strip().strip('"').strip()
        nums = t.split()
        if cur is not None and len(nums) == n and all(x.isdigit() for x in nums):
            cur.append([int(x) for x in nums])
        elif cur is not None and cur and not nums:
            mats.append((cur, aut))
            cur = None
    if cur:
        mats.append((cur, aut))
    return mats


def kfact():
    res = {}
    p10 = os.path.join(REF, "k-Factorization/Results/Complete_graphs/10/10x9x2/K_OrderedMatrices.txt")
    if os.path.exists(p10):
        mats = parse_kfact(read(p10), 10)
        res["k10"] = dict(n=10, source=os.path.relpath(p10, REF),
                          matrices=[dict(rounds=m, aut=a) for m, a in mats])
    for key, n, rel in (("p10", 10, "10/10x9x2/P_OrderedMatrices.txt"), ("p14", 14, "14/14x13x2/P_OrderedMatrices.txt"),
                        ("u8", 8, "8/8x7x2_44/U_OrderedMatrices.txt"), ("u10", 10, "10/10x9x2_46/U_OrderedMatrices.txt"),
                        ("k15x3", 15, "15/15x7x3/K_OrderedMatrices.txt"), ("p15x3", 15, "15/15x7x3/P_OrderedMatrices.txt")):
        path = os.path.join(REF, "k-Factorization/Results/Complete_graphs", rel)
        if os.path.exists(path):
            mats = parse_kfact(read(path), n)
            res[key] = dict(n=n, source=os.path.relpath(path, REF),
                            matrices=[dict(rounds=m, aut=a) for m, a in mats])
    p12 = glob.glob(os.path.join(REF, "k-Factorization/Results/Complete_graphs/12/12x11x2_66/U_OrderedMatrices.txt"))
    if p12:
        mats = parse_kfact(read(p12[0]), 12)
        res["u12"] = dict(n=12, source=os.path.relpath(p12[0], REF),
                          matrices=[dict(rounds=m, aut=a) for m, a in mats])
    return res


def text_blocks():
    """Verbatim output lines (expected text for the output formatters): the
    "BIBD(...)" header through the "Total:" line of the result files written
    in the current table layout, and the first printed matrix blocks of a
    Quasi-Combined result file."""
    tables = {}
    for name in ["V =   7/7_3_1.txt"] + [f"V =   6/6_3_{l}.txt" for l in (2, 4, 6, 8, 10, 12)]:
        path = os.path.join(REF, "Results/BIBDs", name)
        if not os.path.exists(path):
            continue
        lines = read(path).splitlines()
        start = next(i for i, t in enumerate(lines) if t.startswith("BIBD("))
        end = next(i for i, t in enumerate(lines) if t.strip().startswith("Total:"))
        key = os.path.basename(name)[:-4]
        tables[key] = dict(source=os.path.relpath(path, REF), lines=lines[start:end + 1])
    mpath = os.path.join(REF, "Results/Quasi-Combined_BIBDs/V =   9/9_3_2.txt")
    blocks = []
    if os.path.exists(mpath):
        lines = read(mpath).splitlines()
        i = 0
        while i < len(lines) and len(blocks) < 12:
            if lines[i].startswith("Matrix #"):
                j = i + 1
                while j < len(lines) and lines[j] and set(lines[j]) <= set("01"):
                    j += 1
                blocks.append(lines[i:j])
                i = j
            else:
                i += 1
    return dict(tables=tables, matrices=dict(source=os.path.relpath(mpath, REF), blocks=blocks),
                kfact=kfact_blocks())


def kfact_blocks():
    """Verbatim 1-factorization blocks ('<i>: |Aut(M)| = a,  Cycles:...' and the
    quoted rows) of the k-Sys result files: the first 60 of K_10 and every
    block of the perfect and uniform files."""
    out = []
    for n, rel, limit in ((10, "10/10x9x2/K_OrderedMatrices.txt", 60), (10, "10/10x9x2/P_OrderedMatrices.txt", None),
                          (14, "14/14x13x2/P_OrderedMatrices.txt", None), (8, "8/8x7x2_44/U_OrderedMatrices.txt", None),
                          (10, "10/10x9x2_46/U_OrderedMatrices.txt", None),
                          (12, "12/12x11x2_66/U_OrderedMatrices.txt", None)):
        path = os.path.join(REF, "k-Factorization/Results/Complete_graphs", rel)
        if not os.path.exists(path):
            continue
        lines = read(path).splitlines()
        blocks = []
        for i, t in enumerate(lines):
            if re.match(r"\s*\d+: \|Aut\(M\)\| =", t):
                j = i + 1
                while j < len(lines) and lines[j].startswith(' "'):
                    j += 1
                blocks.append(lines[i:j])
                if limit and len(blocks) >= limit:
                    break
        out.append(dict(n=n, source=os.path.relpath(path, REF), blocks=blocks))
    return out


def main():
    counts = bibd_counts()
    with open(os.path.join(OUT, "bibd_counts.json"), "w") as f:
        json.dump(counts, f, indent=0)
    printed = printed_bibds()
    with open(os.path.join(OUT, "printed_bibds.json"), "w") as f:
        json.dump(printed, f, indent=0)
    kf = kfact()
    with open(os.path.join(OUT, "kfact.json"), "w") as f:
        json.dump(kf, f, indent=0)
    blocks = text_blocks()
    with open(os.path.join(OUT, "format_golden.json"), "w") as f:
        json.dump(blocks, f, indent=0)
    print(len(counts), "count tables;", sum(len(p["matrices"]) for p in printed), "printed BIBDs;",
          {k: len(v["matrices"]) for k, v in kf.items()}, "k-fact matrices")


if __name__ == "__main__":
    main()
