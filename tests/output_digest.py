"""Digest of the XDMF + HDF5 output of a run, in the layout of the reference's
writer (io_xdmf.c): the time series <base>_ts.xmf, the global file
<base>_gbl.h5 and the per-rank files <base>_r<rank>/<step>.h5.

HDF5 files are read through the h5dump tool of the image's HDF5 (h5py is not
installed): `h5dump -p -H` gives the structure (datasets, types, shapes,
chunking, filters, fill time and the stored size after compression) and
`h5dump -b LE` the raw data of every dataset, which is hashed per dataset.
The time series is kept as text without its wall-clock date line.

VTK output (io_vtk.c) holds no date, so its digest is a hash of every
<base>_<step>.pvtu and <base>_<rank>/<step>.vtu file."""
import glob
import hashlib
import os
import re
import shutil
import subprocess
import tempfile

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

_TYPE_BYTES = {"H5T_IEEE_F32LE": 4, "H5T_IEEE_F64LE": 8, "H5T_STD_I32LE": 4, "H5T_STD_U32LE": 4,
               "H5T_STD_U64LE": 8, "H5T_STD_I64LE": 8}


def available():
    return os.path.exists(H5DUMP)


def _datasets(header):
    """(path, bytes) of every dataset in the order h5dump lists (and dumps) them."""
    out, groups, name, nbytes, depth = [], [], None, 0, 0
    for line in header.splitlines():
        s = line.strip()
        m = re.match(r'GROUP "(.*)" \{', s)
        if m:
            groups.append(m.group(1).strip("/"))
            continue
        m = re.match(r'DATASET "(.*)" \{', s)
        if m:
            name, depth = m.group(1), 1
            continue
        if name is not None:
            if s.startswith("DATATYPE"):
                nbytes = _TYPE_BYTES[s.split()[1]]
            m = re.match(r"DATASPACE\s+SIMPLE \{ \( ([0-9, ]+) \)", s)
            if m:
                for d in m.group(1).split(","):
                    nbytes *= int(d)
            depth += s.count("{") - s.count("}")
            if depth == 0:
                path = "/".join(g for g in groups if g) + "/" + name
                out.append(("/" + path.lstrip("/"), nbytes))
                name = None
            continue
        if s == "}" and groups:
            groups.pop()
    return out


def h5_digest(path):
    """{"header": structure text, "data": {dataset: sha256[:16]}} of one file."""
    header = subprocess.run([H5DUMP, "-p", "-H", path], stdout=subprocess.PIPE, check=True).stdout.decode()
    lines = header.splitlines()
    lines[0] = "HDF5 {"                       # the first line names the file
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "data.bin")
        subprocess.run([H5DUMP, "-b", "LE", "-o", raw, path], stdout=subprocess.DEVNULL, check=True)
        blob = open(raw, "rb").read()
    data, off = {}, 0
    for ds, n in _datasets(header):
        data[ds] = hashlib.sha256(blob[off:off + n]).hexdigest()[:16]
        off += n
    assert off == len(blob), (path, off, len(blob))
    return {"header": "\n".join(lines), "data": data}


def xmf_text(path):
    """The time series without the wall-clock line under "Simulation Date"."""
    lines = open(path).read().splitlines()
    for i, l in enumerate(lines):
        if 'Name="Simulation Date"' in l:
            del lines[i + 1]
            break
    return "\n".join(lines) + "\n"


def digest_output(root, base):
    """Digest of a run's output under directory root for BaseFileName base."""
    files = {}
    prefix = os.path.join(root, base)
    steps = sorted(glob.glob(prefix + "_r*/[0-9][0-9][0-9][0-9][0-9].h5"))   # not checkpoint files
    for p in steps + [prefix + "_gbl.h5"]:
        files[os.path.relpath(p, root)] = h5_digest(p)
    return {"xmf": xmf_text(prefix + "_ts.xmf"), "files": files}


def compare(got, want):
    """List of differences between two digests (empty when they agree)."""
    bad = []
    if got["xmf"] != want["xmf"]:
        a, b = got["xmf"].splitlines(), want["xmf"].splitlines()
        first = next((i for i, (x, y) in enumerate(zip(a, b)) if x != y), min(len(a), len(b)))
        bad.append("xmf differs from line %d (%d vs %d lines)" % (first + 1, len(a), len(b)))
    for f in sorted(set(got["files"]) | set(want["files"])):
        if f not in got["files"] or f not in want["files"]:
            bad.append("%s only in %s" % (f, "output" if f in got["files"] else "fixture"))
            continue
        g, w = got["files"][f], want["files"][f]
        if g["header"] != w["header"]:
            bad.append("%s: structure differs" % f)
        for ds in sorted(set(g["data"]) | set(w["data"])):
            if g["data"].get(ds) != w["data"].get(ds):
                bad.append("%s:%s data differs" % (f, ds))
    return bad


def vtk_digest(root, base):
    """{relative path: sha256[:16]} of a run's VTK files under root."""
    prefix = os.path.join(root, base)
    paths = sorted(glob.glob(prefix + "_*.pvtu") + glob.glob(prefix + "_[0-9][0-9][0-9][0-9]/*.vtu"))
    return {os.path.relpath(p, root): hashlib.sha256(open(p, "rb").read()).hexdigest()[:16] for p in paths}


def _dataset_paths(header):
    """Paths of every dataset in an `h5dump -H` listing, compound types included."""
    out, stack = [], []
    for line in header.splitlines():
        s = line.strip()
        m = re.match(r'(GROUP|DATASET|DATATYPE|DATASPACE|STORAGE_LAYOUT|FILTERS|FILLVALUE|ALLOCATION_TIME)\b'
                     r'(?: "(.*)")?.*\{\s*$', s)
        if m:
            kind, name = m.group(1), m.group(2)
            if kind == "DATASET":
                out.append("/" + "/".join(p for p in [n for k, n in stack if k == "GROUP"] + [name] if p))
            stack.append((kind, (name or "").strip("/")))
            continue
        opened, closed = s.count("{"), s.count("}")
        for _ in range(opened):
            stack.append(("other", ""))
        for _ in range(closed):
            if stack:
                stack.pop()
    return out


def _drop_group(lines, name):
    """The listing without the top-level group `name` and its contents."""
    out, depth = [], 0
    for line in lines:
        s = line.strip()
        if depth == 0 and s == 'GROUP "%s" {' % name:
            depth = 1
            continue
        if depth > 0:
            depth += s.count("{") - s.count("}")
            continue
        out.append(line)
    return out


def h5_digest_each(path, drop=()):
    """As h5_digest, with every dataset dumped on its own (compound types
    have no fixed entry in the size table); top-level groups named in
    `drop` are left out."""
    header = subprocess.run([H5DUMP, "-p", "-H", path], stdout=subprocess.PIPE, check=True).stdout.decode()
    lines = header.splitlines()
    lines[0] = "HDF5 {"
    for g in drop:
        lines = _drop_group(lines, g)
    header = "\n".join(lines)
    data = {}
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "data.bin")
        for ds in _dataset_paths(header):
            r = subprocess.run([H5DUMP, "-b", "LE", "-d", ds, "-o", raw, path], stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL)
            if r.returncode == 0:
                blob = open(raw, "rb").read()
            else:
                # h5dump does not write nested compound members in binary:
                # their values as text, every double to 17 digits
                txt = subprocess.run([H5DUMP, "-m", "%.17g", "-d", ds, path], stdout=subprocess.PIPE,
                                     check=True).stdout.decode()
                blob = txt[txt.index("DATA {"):].encode()
            data[ds] = hashlib.sha256(blob).hexdigest()[:16]
    return {"header": "\n".join(lines), "data": data}


def checkpoint_digest(root, drop=()):
    """Digest of the checkpoint files a run left in root: cp_main.h5 and
    cp_task_<rank>.h5 (writeTaskCheckpoint's layout, checkpoint.c:955-1238),
    without the top-level groups named in `drop`."""
    files = sorted(glob.glob(os.path.join(root, "cp_main.h5")) + glob.glob(os.path.join(root, "cp_task_*.h5")))
    return {os.path.basename(p): h5_digest_each(p, drop) for p in files}


def compare_files(got, want, skip=()):
    """Differences between two {file: digest} maps; datasets whose path ends
    in one of `skip` are left out."""
    bad = []
    for f in sorted(set(got) | set(want)):
        if f not in got or f not in want:
            bad.append("%s only in %s" % (f, "output" if f in got else "fixture"))
            continue
        g, w = got[f], want[f]
        if g["header"] != w["header"]:
            a, b = g["header"].splitlines(), w["header"].splitlines()
            first = next((i for i, (x, y) in enumerate(zip(a, b)) if x != y), min(len(a), len(b)))
            bad.append("%s: structure differs from line %d: %r vs %r" % (
                f, first + 1, a[first] if first < len(a) else None, b[first] if first < len(b) else None))
        for ds in sorted(set(g["data"]) | set(w["data"])):
            if ds.endswith(tuple(skip)):
                continue
            if g["data"].get(ds) != w["data"].get(ds):
                bad.append("%s:%s data differs" % (f, ds))
    return bad
