"""Pure-Python restatement of the text edge-list ingest (the checker of
gala_ingest_text and gala_ingest_text_device).

Test infrastructure only: imported by tests/, never by the product path.
It follows Graph::Graph(file, is_wt), LinXi-lx/GALA-mg src/graph/graph.cpp:
  :27-51   one line at a time, empty and '#' lines skipped, both directions
           pushed into per-source rows (edges[src], edges[dest]);
  :54-75   per-row std::map: entries in ascending neighbour order, the weight
           of a repeated neighbour added to the first one's;
  :77-103  linked vertices renumbered densely in ascending original id;
  :111-126 inclusive row ends and the neighbour array.
The reference leaves malformed lines undefined (it reads uninitialised ids);
the rules for them and for weight tokens are the library's own, restated
here from its header: two leading integer tokens make an edge line, other
lines are skipped; ids outside [0, 2^31-1] and malformed weights are errors;
a float weight is the decimal rounded to the nearest double (Python's float()
of the token, correctly rounded), sums are taken in line order in float64 and
rounded once to float32; int weights are summed exactly and must fit int32.
The reference's own defects (a self-loop line truncates the neighbour array;
the weighted writer stores only weights[0]) are not restated: a self-loop is
one entry whose weight is pushed twice.
Small inputs only (pure-Python loops).
"""
import re

import numpy as np

W_NONE, W_I32, W_F32 = 0, 1, 2

_WS = " \t\r\v\f"
_INT = re.compile(r"[+-]?[0-9]+\Z")
_DEC = re.compile(r"[+-]?([0-9]+(\.[0-9]*)?|\.[0-9]+)([eE][+-]?[0-9]+)?\Z")


class IngestError(ValueError):
    pass


def _tokens(line):
    toks, cur = [], []
    for ch in line:
        if ch in _WS:
            if cur:
                toks.append("".join(cur))
                cur = []
        else:
            cur.append(ch)
    if cur:
        toks.append("".join(cur))
    return toks


def _sig_digits(tok):
    mant = re.split(r"[eE]", tok.lstrip("+-"))[0].replace(".", "")
    core = mant.strip("0")
    return len(core)


def parse_lines(data, weight_mode):
    """[(a, b, w)] edge lines of `data` (bytes) in file order."""
    text = data.decode("latin-1")
    out = []
    for line in text.split("\n"):  # graph.cpp:27 getline
        if not line or line[0] == "#":  # graph.cpp:34
            continue
        t = _tokens(line)
        if len(t) < 2 or not _INT.match(t[0]) or not _INT.match(t[1]):
            continue
        a, b = int(t[0]), int(t[1])
        if not (0 <= a <= 2**31 - 1 and 0 <= b <= 2**31 - 1):
            raise IngestError("vertex id out of range")
        w = 1
        if weight_mode == W_I32:
            if len(t) > 2:
                if not _INT.match(t[2]) or not (-(2**31) <= int(t[2]) <= 2**31 - 1):
                    raise IngestError("bad int weight")
                w = int(t[2])
        elif weight_mode == W_F32:
            w = 1.0
            if len(t) > 2:
                if not _DEC.match(t[2]) or _sig_digits(t[2]) > 40:
                    raise IngestError("bad decimal weight")
                w = float(t[2])
                if w in (float("inf"), float("-inf")):
                    raise IngestError("weight out of range")
        out.append((a, b, w))
    return out


def ingest(data, weight_mode=W_NONE):
    """(n, end_incl int64[n], nbr int32[m], w or None) for the edge list `data`."""
    edges = parse_lines(data, weight_mode)
    rows = {}  # original id -> {neighbour original id: accumulated weight}, insertion = line order
    for a, b, w in edges:  # graph.cpp:45-47: both directions
        for src, dst in ((a, b), (b, a)):
            r = rows.setdefault(src, {})
            if dst in r:
                r[dst] = r[dst] + w  # graph.cpp:63-64 (weighted: summed)
            else:
                r[dst] = w
    ids = sorted(rows)  # linked vertices; every one owns a row
    renum = {v: i for i, v in enumerate(ids)}
    end, nbr, wts = [], [], []
    for v in ids:
        r = rows[v]
        for u in sorted(r, key=lambda x: renum[x]):
            nbr.append(renum[u])
            wts.append(r[u])
        end.append(len(nbr))
    w = None
    if weight_mode == W_F32:
        with np.errstate(over="ignore"):
            w = np.array(wts, np.float64).astype(np.float32)
        if not np.all(np.isfinite(w)):
            raise IngestError("weight sum out of float32 range")
    elif weight_mode == W_I32:
        if any(not (-(2**31) <= x <= 2**31 - 1) for x in wts):
            raise IngestError("weight sum out of int32 range")
        w = np.array(wts, np.int64).astype(np.int32)
    return len(ids), np.array(end, np.int64), np.array(nbr, np.int32), w
