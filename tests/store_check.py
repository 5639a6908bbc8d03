"""Static check of the generated gfx950 code: no instruction overwrites the
data VGPRs of a vector-memory store before s_waitcnt expcnt(0) (or vmcnt(0))
has reported them read.

A store reads its data registers some cycles after it issues; an
overwrite scheduled right behind it can reach memory instead of the stored
value.  This produced sporadic wrong 64-byte groups in the operator's FP64
output (repeated DSq and multi-shift CG runs at U(3) 8^3x16 were not
bitwise repeatable; a wait after each store made them so).  Every
source of the library stores through helpers that hold the data until the
wait (bstore + hold, st_keep, ST); this module compiles each source to device
assembly with the build's flags and lists the kernels that still overwrite
store data.  The scan is linear through each kernel: a store's registers stay
"live" across labels and branches until a full wait, so an overwrite on any
later path is reported.

    python tests/store_check.py [NCOL]    (prints the flagged kernels)
"""
import collections
import concurrent.futures as cf
import os
import re
import subprocess
import sys
import tempfile

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from susy_amd import build as B  # noqa: E402

CHECKED = tuple(B.SOURCES)
FUNC = re.compile(r"^(_Z\S+):\s")
STORE = re.compile(r"^\s*(buffer_store_\w+|global_store_\w+|flat_store_\w+)\s+(v\[\d+:\d+\]|v\d+),"
                   r"\s*(v\[\d+:\d+\]|v\d+)")
WRITE = re.compile(r"^\s*(v_\w+|buffer_load\w*|global_load\w*|flat_load\w*|ds_read\w*|"
                   r"ds_bpermute\w*|ds_swizzle\w*)\s+(v\[\d+:\d+\]|v\d+)")
FULL_WAIT = re.compile(r"^\s*s_waitcnt\b.*(expcnt\(0\)|vmcnt\(0\))")
BLOCK_END = re.compile(r"^\s*s_endpgm")


def _regs(tok):
    m = re.match(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"v(\d+)", tok)
    return {int(m.group(1))} if m else set()


def overwrites(lines):
    """[(kernel, store line, overwriting line)] in device assembly text"""
    found, kernel, live = [], None, []
    for i, line in enumerate(lines):
        m = FUNC.match(line)
        if m:
            kernel, live = m.group(1), []
            continue
        if kernel is None:
            continue
        if FULL_WAIT.match(line) or BLOCK_END.match(line):
            live = []
            if "s_endpgm" in line:
                kernel = None
            continue
        st = STORE.match(line)
        wr = None if st else WRITE.match(line)
        if wr:
            w = _regs(wr.group(2))
            for r, j in live:
                if r & w:
                    found.append((kernel, j, i))
            live = [(r, j) for r, j in live if not (r & w)]
        if st:
            # buffer_store vdata, vaddr, ...; global_store vaddr, vdata, ...
            data = st.group(2) if st.group(1).startswith("buffer") else st.group(3)
            live.append((_regs(data), i))
    return found


def _asm(src, ncol, out):
    cmd = [B.HIPCC, *B.CFLAGS, f"-DNCOL={ncol}", "--cuda-device-only", "-S",
           os.path.join(B.CSRC, src), "-o", out]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("device assembly failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return out


def check(ncol=4, sources=CHECKED):
    """{kernel: number of store-data overwrites}"""
    res = collections.OrderedDict()
    with tempfile.TemporaryDirectory() as tmp:
        outs = [os.path.join(tmp, os.path.splitext(s)[0] + ".s") for s in sources]
        with cf.ThreadPoolExecutor(max_workers=len(sources)) as ex:
            list(ex.map(lambda a: _asm(a[0], ncol, a[1]), zip(sources, outs)))
        for out in outs:
            with open(out) as f:
                for kernel, _, _ in overwrites(f.read().split("\n")):
                    res[kernel] = res.get(kernel, 0) + 1
    return res


if __name__ == "__main__":
    found = check(int(sys.argv[1]) if len(sys.argv) > 1 else 4)
    for k, v in found.items():
        print(f"{v:5d}  {k}")
    print(f"kernels overwriting store data: {len(found)}")
    sys.exit(1 if found else 0)
