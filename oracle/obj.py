"""ORACLE: test infrastructure.  Only tests/ may use it, as the checker of
the library's OBJ/MTL loader (tr_load_obj_rt_geometry).

A second, independent restatement of LoadObjFileStaticRTGeometry
(src/mesh/obj-loader.cc:635-840) in Python: the flex rules of
src/mesh/obj-loader.l and obj-mtl-loader.l as regular expressions with
longest-match / earliest-rule selection, the parsers
(src/mesh/obj-loader-parser.cc:322-515, obj-mtl-parser.cc:148-247), the
driver's index and material handling (obj-loader-driver.hh:51-83,
obj-loader-driver.cc:71-86, obj-mtl-loader-driver.hh:48-54) and the
per-group InterleaveInterm (obj-loader.cc:61-384) with InterleaveVertices
(src/utils/interleave-vertices.cc:62-115), all FP32 arithmetic in numpy
float32 scalars, in the reference's operation order.

Returns plain data: dict(submeshes=[(material index or None, vertex_count,
vertex_offset, base_index, stride)], vertex_data=bytes, index_data=int32
array, materials=[dict(model, diffuse, specular, specular_power)]).
Raises ObjError where the reference logs a diagnostic (and the loader
fails), and where its behaviour is undefined (indices out of range, mixed
vertex formats in a group).  Names after usemtl / mtllib / map_* are
stripped of surrounding blanks and carriage returns, as the library does.
"""
import math
import os
import re

import numpy as np

F = np.float32


class ObjError(Exception):
    pass


_NUM = re.compile(r"-?[0-9]+(\.[0-9]+([eE][+-]?[0-9]*)?|[eE][+-]?[0-9]*)?")
_NAME = re.compile(r"[A-Za-z0-9_.]+")

OBJ_KEYWORDS = ["vp", "vn", "vt", "v", "deg", "bmat", "step", "cstype", "parm", "trim", "hole", "scrv", "sp", "end",
                "con", "p", "l", "f", "curv", "curv2", "surf", "g", "s", "mg", "o", "bevel", "c_interp", "d_interp",
                "lod", "usemtl ", "mtllib ", "shadow_obj", "trace_obj", "ctech", "stech", "off"]
MTL_KEYWORDS = ["newmtl", "Ka", "Kd", "Ks", "Tf", "illum", "Ns", "sharpness", "Ni", "map_Ka ", "map_Kd ", "map_Ks ",
                "map_Ns ", "map_d ", "d", "Tr", "Ke"]


def tokenize(text, keywords, mtl=False):
    """List of (kind, value, line): kind in eol, int, float, str, '/', kw,
    illum (the MTL IllumModel token, value = character code of the digit).
    A keyword ending in a blank stands for the rule `word[ \\t]` and is
    followed by the rest of the line as one str token."""
    toks, i, line, n = [], 0, 1, len(text)
    while i < n:
        c = text[i]
        if c in " \t":
            i += 1
            continue
        if c == "#":
            while i < n and text[i] not in "\r\n":
                i += 1
            continue
        if c in "\r\n":
            j = i
            while j < n and text[j] == "\r":
                j += 1
            if j < n and text[j] == "\n":
                toks.append(("eol", None, line))
                line += 1
                i = j + 1
                continue
            raise ObjError("line %d: invalid character" % line)
        if c == "/":
            toks.append(("/", None, line))
            i += 1
            continue
        cands = []  # (length, -priority, kind, value)
        m = _NUM.match(text, i)
        if m:
            s = m.group(0)
            kind = "float" if ("." in s or "e" in s or "E" in s) else "int"
            cands.append((len(s), 3, kind, s))
        for kw in keywords:
            word = kw.rstrip(" ")
            if text.startswith(word, i):
                ln = len(word)
                if kw.endswith(" "):
                    if i + ln >= n or text[i + ln] not in " \t":
                        continue
                    ln += 1
                cands.append((ln, 2, "kw", kw))
        if mtl and text.startswith("illum_", i) and i + 6 < n and text[i + 6].isdigit():
            cands.append((7, 2, "illum", ord(text[i + 6])))
        m = _NAME.match(text, i)
        if m:
            cands.append((len(m.group(0)), 1, "str", m.group(0)))
        if not cands:
            raise ObjError("line %d: invalid character" % line)
        ln, _, kind, val = max(cands, key=lambda t: (t[0], t[1]))
        i += ln
        if kind == "int":
            val = int(val)
        elif kind == "float":
            val = F(float(val)) if not re.search(r"[eE][+-]?$", val) else F(float(re.sub(r"[eE][+-]?$", "", val)))
        toks.append((kind, val, line))
        if kind == "kw" and val.endswith(" "):
            # rest-of-line name state: [\r]*\n ends it, [^\n]* is the name
            j = i
            while j < n and text[j] != "\n":
                j += 1
            r = i
            while r < n and text[r] == "\r":
                r += 1
            eol_len = (r - i + 1) if (r < n and text[r] == "\n") else 0
            if eol_len > j - i:
                toks.append(("eol", None, line))
                line += 1
                i += eol_len
            elif j > i:
                toks.append(("str", text[i:j], line))
                i = j
                if i < n:  # the newline itself
                    toks.append(("eol", None, line))
                    line += 1
                    i += 1
    return toks


class _Reader:
    def __init__(self, toks, name):
        self.t, self.k, self.name = toks, 0, name

    def next(self):
        if self.k < len(self.t):
            self.k += 1
            return self.t[self.k - 1]
        return ("eof", None, -1)

    def fail(self, what, tok=None):
        raise ObjError("%s: %s" % (self.name, what))

    def number(self):
        kind, v, _ = self.next()
        if kind == "float":
            return F(v)
        if kind == "int":
            return F(v)
        self.fail("expecting a number")

    def index(self):
        kind, v, _ = self.next()
        if kind != "int":
            self.fail("expecting an index")
        return v

    def slash(self):
        if self.next()[0] != "/":
            self.fail("expecting /")

    def end(self):
        if self.next()[0] not in ("eol", "eof"):
            self.fail("expecting end of statement")


def _strip(s):
    return s.strip(" \t\r")


def parse_mtl(path):
    """Materials of an MTL file as dicts with the ObjMtlLoader::Material
    defaults (obj-mtl-loader.hh:57-76)."""
    try:
        with open(path, "rb") as f:
            text = f.read().decode("latin-1")
    except OSError:
        raise ObjError("cannot read material file " + path)
    rd = _Reader(tokenize(text, MTL_KEYWORDS, mtl=True), path)
    mats, cur = [], None
    v3 = {"Ka": "ambient", "Kd": "diffuse", "Ks": "specular", "Ke": "emission", "Tf": "transmission"}
    while True:
        kind, v, _ = rd.next()
        if kind == "eof":
            break
        if kind == "eol":
            continue
        if kind != "kw":
            rd.fail("unexpected top-level token")
        if v == "newmtl":
            k2, name, _ = rd.next()
            if k2 != "str":
                rd.fail("expecting a material name")
            cur = dict(name=name.lower(), emission=(F(0), F(0), F(0)), diffuse=(F(1), F(1), F(1)),
                       specular=(F(0), F(0), F(0)), illum=0, ns=F(1), map_kd="")
            mats.append(cur)
            continue
        if v in v3:
            x, y, z = rd.number(), rd.number(), rd.number()
            rd.end()
            if cur is None:
                rd.fail("no material")
            cur[v3[v]] = (x, y, z)
        elif v == "illum":
            k2, num, _ = rd.next()
            if k2 not in ("int", "illum"):
                rd.fail("invalid illumination model")
            rd.end()
            if cur is None:
                rd.fail("no material")
            cur["illum"] = num
        elif v in ("d", "Tr"):
            rd.number()
            k2 = rd.next()[0]
            if k2 in ("int", "float"):
                k2 = rd.next()[0]
            if k2 not in ("eol", "eof"):
                rd.fail("expecting end of statement")
            if cur is None:
                rd.fail("no material")
        elif v in ("Ns", "sharpness", "Ni"):
            x = rd.number()
            rd.end()
            if cur is None:
                rd.fail("no material")
            if v == "Ns":
                cur["ns"] = x
        else:  # map_*
            k2, name, _ = rd.next()
            if k2 != "str":
                rd.fail("expecting a file name")
            rd.end()
            if cur is None:
                rd.fail("no material")
            if v == "map_Kd ":
                cur["map_kd"] = _strip(name)
    return mats


def _norm_self(v):
    """NormalizeSelf (vector3.hh:78-88)."""
    x, y, z = v
    ln = F(np.sqrt(F(F(x * x + y * y) + z * z)))
    if ln != F(0):
        r = F(F(1) / ln)
        return (F(x * r), F(y * r), F(z * r))
    return v


def _normalize(v):
    """Normalize (vector3.hh:215-219); Vector3 / float multiplies by the
    reciprocal (vector3.hh:185-190)."""
    x, y, z = v
    ln = F(np.sqrt(F(F(x * x + y * y) + z * z)))
    if ln == F(0):
        return v
    r = F(F(1) / ln)
    return (F(x * r), F(y * r), F(z * r))


def _sub(a, b):
    return (F(a[0] - b[0]), F(a[1] - b[1]), F(a[2] - b[2]))


def _add(a, b):
    return (F(a[0] + b[0]), F(a[1] + b[1]), F(a[2] + b[2]))


def _cross(a, b):
    return (F(a[1] * b[2] - a[2] * b[1]), F(a[2] * b[0] - a[0] * b[2]), F(a[0] * b[1] - a[1] * b[0]))


def _dot(a, b):
    return F(F(a[0] * b[0] + a[1] * b[1]) + a[2] * b[2])


def _tangent(p0, p1, p2, t0, t1, t2):
    """GenerateTangentSpace (math/triangle.hh:52-92)."""
    e0, e1 = _sub(p1, p0), _sub(p2, p0)
    c0 = (F(t1[0] - t0[0]), F(t1[1] - t0[1]))
    c1 = (F(t2[0] - t0[0]), F(t2[1] - t0[1]))
    with np.errstate(all="ignore"):
        det = F(c0[0] * c1[1] - c0[1] * c1[0])
        i0 = (F(c1[1] / det), F(-c0[1] / det))
        i1 = (F(-c1[0] / det), F(c0[0] / det))
    if not np.isfinite(i0[0]):
        n = _cross(e0, e1)
        s = _add(e0, e1)
        if c0[0] != F(0):
            sg = F(1) if c0[0] >= F(0) else F(-1)
            return (F(s[0] * sg), F(s[1] * sg), F(s[2] * sg))
        sg = F(1) if c0[1] >= F(0) else F(-1)
        return _cross((F(s[0] * sg), F(s[1] * sg), F(s[2] * sg)), n)

    def tri(a, b):
        return (F(a * i0[0] + b * i0[1]), F(a * i1[0] + b * i1[1]))
    x, y, z = tri(e0[0], e1[0]), tri(e0[1], e1[1]), tri(e0[2], e1[2])
    return (x[0], y[0], z[0])


def _face(rd, d):
    def push(p, t, n):
        d["pi"].append(p - 1 if p >= 0 else len(d["pos"]) + p)
        if t is not None:
            d["ti"].append(t - 1 if t >= 0 else len(d["tc"]) + t)
        if n is not None:
            d["ni"].append(n - 1 if n >= 0 else len(d["norm"]) + n)

    verts = []  # (p, t, n) per face vertex, then fan
    sp = rd.index()
    kind, v, _ = rd.next()
    if kind == "int":
        verts = [(sp, None, None), (v, None, None), (rd.index(), None, None)]
        while True:
            kind, v, _ = rd.next()
            if kind != "int":
                break
            verts.append((v, None, None))
    elif kind == "/":
        kind, v, _ = rd.next()
        if kind == "int":
            st = v
            kind, v, _ = rd.next()
            if kind == "int":
                rd.slash()
                verts = [(sp, st, None), (v, rd.index(), None)]
                cp = rd.index()
                while True:
                    rd.slash()
                    verts.append((cp, rd.index(), None))
                    kind, cp, _ = rd.next()
                    if kind != "int":
                        break
            elif kind == "/":
                sn = rd.index()
                pp = rd.index()
                rd.slash()
                pt = rd.index()
                rd.slash()
                verts = [(sp, st, sn), (pp, pt, rd.index())]
                cp = rd.index()
                while True:
                    rd.slash()
                    ct = rd.index()
                    rd.slash()
                    verts.append((cp, ct, rd.index()))
                    kind, cp, _ = rd.next()
                    if kind != "int":
                        break
            else:
                rd.fail("invalid face")
        elif kind == "/":
            sn = rd.index()
            pp = rd.index()
            rd.slash()
            rd.slash()
            verts = [(sp, None, sn), (pp, None, rd.index())]
            cp = rd.index()
            while True:
                rd.slash()
                rd.slash()
                verts.append((cp, None, rd.index()))
                kind, cp, _ = rd.next()
                if kind != "int":
                    break
        else:
            rd.fail("invalid face")
    else:
        rd.fail("invalid face")
    if kind not in ("eol", "eof"):
        rd.fail("invalid face")
    for k in range(1, len(verts) - 1):
        for vv in (verts[0], verts[k], verts[k + 1]):
            push(*vv)


def parse_obj(path):
    try:
        with open(path, "rb") as f:
            text = f.read().decode("latin-1")
    except OSError:
        raise ObjError("cannot read object file " + path)
    rd = _Reader(tokenize(text, OBJ_KEYWORDS), path)
    d = dict(pos=[], tc=[], norm=[], pi=[], ti=[], ni=[], groups=[], mats=[], cur=0)
    base = os.path.dirname(path) or "."
    while True:
        kind, v, _ = rd.next()
        if kind == "eof":
            break
        if kind == "eol":
            continue
        if kind != "kw":
            rd.fail("unexpected top-level token")
        if v in ("p", "l", "s"):
            while rd.next()[0] not in ("eol", "eof"):
                pass
        elif v == "v":
            x, y, z = rd.number(), rd.number(), rd.number()
            w = F(1)
            k2, v2, _ = rd.next()
            if k2 in ("int", "float"):
                w = F(v2)
            elif k2 not in ("eol", "eof"):
                rd.fail("bad vertex")
            f = F(F(1) / w)
            d["pos"].append((F(x * f), F(y * f), F(z * f)))
        elif v == "vn":
            x, y, z = rd.number(), rd.number(), rd.number()
            rd.end()
            d["norm"].append((x, y, z))
        elif v == "vt":
            x, y = rd.number(), rd.number()
            k2 = rd.next()[0]
            if k2 in ("int", "float"):
                k2 = rd.next()[0]
            if k2 not in ("eol", "eof"):
                rd.fail("bad texcoord")
            d["tc"].append((x, y))
        elif v in ("g", "o"):
            if rd.next()[0] != "str":
                rd.fail("bad group")
            d["groups"].append([len(d["pi"]), len(d["ti"]), len(d["ni"]), d["cur"]])
        elif v == "f":
            _face(rd, d)
        elif v == "usemtl ":
            k2, name, _ = rd.next()
            if k2 != "str":
                rd.fail("bad usemtl")
            name = _strip(name).lower()
            idx = [m["name"] for m in d["mats"]]
            if name not in idx:
                rd.fail("unknown material " + name)
            d["cur"] = idx.index(name)
            if d["groups"]:
                d["groups"][-1][3] = d["cur"]
        elif v == "mtllib ":
            k2, name, _ = rd.next()
            if k2 != "str":
                rd.fail("bad mtllib")
            d["mats"] += parse_mtl(os.path.join(base, _strip(name)))
        else:
            rd.fail("unsupported statement " + v)
    return d


def _interleave(d, g, pn, tn, nn, flags, data, inds):
    if pn == 0:
        return 0
    if pn % 3 or (tn and tn != pn) or (nn and nn != pn):
        raise ObjError("mixed vertex formats")
    pi = d["pi"][g[0]:g[0] + pn]
    ti = d["ti"][g[1]:g[1] + tn] if tn else None
    ni = d["ni"][g[2]:g[2] + nn] if nn else None
    for arr, src in ((pi, d["pos"]), (ti, d["tc"]), (ni, d["norm"])):
        if arr is not None and any(i < 0 or i >= len(src) for i in arr):
            raise ObjError("index out of range")
    gen_tan = bool(flags & 1) and tn != 0
    consistent = bool(flags & 2)
    mn = min(pi)
    gi = [p - mn for p in pi]
    arrays = [([d["pos"][p] for p in pi], 12)]
    if tn:
        arrays.append(([d["tc"][t] for t in ti], 8))
    if gen_tan:
        tans = {}
        for k in range(0, pn, 3):
            t = _norm_self(_tangent(*(d["pos"][p] for p in pi[k:k + 3]), *(d["tc"][t] for t in ti[k:k + 3])))
            for j in range(3):
                tans[gi[k + j]] = _add(tans.get(gi[k + j], (F(0), F(0), F(0))), t)
        tans = {k: _norm_self(v) for k, v in tans.items()}
        arrays.append(([tans[i] for i in gi], 12))
    if nn:
        norms, nidx = d["norm"], ni
    else:
        acc = {}
        for k in range(0, pn, 3):
            prev, cur, nxt = (d["pos"][p] for p in pi[k:k + 3])
            n = _cross(_sub(nxt, cur), _sub(prev, cur))
            for j in range(3):
                acc[gi[k + j]] = _add(acc.get(gi[k + j], (F(0), F(0), F(0))), n)
        norms = [_norm_self(acc.get(i, (F(0), F(0), F(0)))) for i in range(max(gi) + 1)]
        nidx = gi
    if consistent:
        nmin = min(nidx)
        coef = {}
        for k in range(0, pn, 3):
            a, b, c = (norms[i] for i in nidx[k:k + 3])
            low = min(_dot(a, b), _dot(c, b))
            low = min(low, _dot(a, c)) if _dot(a, c) < low else low
            for j in range(3):
                key = nidx[k + j] - nmin
                coef[key] = min(coef.get(key, F(1)), low) if low < coef.get(key, F(1)) else coef.get(key, F(1))
        out = {}
        for key, cv in coef.items():
            om = F(F(1) - cv)
            out[key] = F(F(np.arccos(cv)) * F(F(1) + F(F(F(0.03632) * om) * om)))
        arrays.append(([out[i - nmin] for i in nidx], 4))
    arrays.append(([norms[i] for i in nidx], 12))
    stride = sum(sz for _, sz in arrays)
    keys = [tuple(x) for x in zip(pi, ti if tn else [0] * pn, nidx)]
    seen = {}
    start = len(data)
    rows = []
    for k, key in enumerate(keys):
        if key in seen:
            inds.append(seen[key])
            continue
        seen[key] = len(seen)
        inds.append(seen[key])
        rows.append(k)
    for k in rows:
        rec = b""
        for vals, sz in arrays:
            rec += np.asarray(vals[k], dtype=np.float32).tobytes()
        data += rec
    if gen_tan:
        for o in range(start, len(data), stride):
            t = tuple(np.frombuffer(bytes(data[o + 20:o + 32]), dtype=np.float32))
            n = tuple(np.frombuffer(bytes(data[o + stride - 12:o + stride]), dtype=np.float32))
            dn = _dot(n, t)
            t = _normalize(_sub(t, (F(n[0] * dn), F(n[1] * dn), F(n[2] * dn))))
            data[o + 20:o + 32] = np.asarray(t, dtype=np.float32).tobytes()
    return stride


def load(path, flags=0):
    d = parse_obj(path)
    data, inds, subs = bytearray(), [], []
    groups = d["groups"]
    for k, g in enumerate(groups):
        nxt = groups[k + 1] if k + 1 < len(groups) else [len(d["pi"]), len(d["ti"]), len(d["ni"])]
        prev_i, prev_v = len(inds), len(data)
        stride = _interleave(d, g, nxt[0] - g[0], nxt[1] - g[1], nxt[2] - g[2], flags, data, inds)
        mat = g[3] if g[3] < len(d["mats"]) else None
        subs.append((mat, len(inds) - prev_i, prev_v, prev_i, stride))
    mats = []
    for m in d["mats"]:
        diffuse = list(m["diffuse"])
        if m["illum"] == 2:
            spec, power = list(m["specular"]), (m["ns"], m["ns"])
        else:
            spec, power = [F(0)] * 3, (F(1), F(1))
        model = 1
        if all(c != F(0) for c in m["emission"]):
            model, diffuse = 0, list(m["emission"])
        if model == 1:  # DefaultMicrofacetMaterialSetup (illumination-models.hh:599-614)
            mx = max(F(a + b) for a, b in zip(diffuse, spec))
            if mx > F(1):  # Spectrum /= max: times the reciprocal (vector3.hh:194-200)
                r = F(F(1) / mx)
                diffuse = [F(a * r) for a in diffuse]
                spec = [F(a * r) for a in spec]
        mats.append(dict(model=model, diffuse=tuple(diffuse), specular=tuple(spec), specular_power=power,
                         map_kd=m["map_kd"]))
    return dict(submeshes=subs, vertex_data=bytes(data), index_data=np.asarray(inds, dtype=np.int32),
                materials=mats)
