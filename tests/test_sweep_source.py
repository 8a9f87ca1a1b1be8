"""Circuit-specialised sweep sources (qlx_plan_sweep_source): what the windowed executor compiles
with hipRTC when qlx_exec_opts::specialize is set. CPU-only: the generated source is checked for
its structure (one load and one store per register, constants in place of descriptors, CX and
SWAP as register renaming) and compiled for gfx950 with hipRTC, which needs no GPU."""
import ctypes as C
import ctypes.util
import os
import re

import numpy as np
import pytest

import qlx

HIPRTC = "/opt/rocm/lib/libhiprtc.so"


def compile_ok(src):
    rtc = C.CDLL(HIPRTC)
    prog = C.c_void_p()
    assert rtc.hiprtcCreateProgram(C.byref(prog), src.encode(), b"s.hip", 0, None, None) == 0
    opts = (C.c_char_p * 6)(b"--offload-arch=gfx950", b"-O3", b"-Xclang", b"-target-feature", b"-Xclang",
                           b"-packed-fp32-ops")
    rc = rtc.hiprtcCompileProgram(prog, 6, opts)
    log = b""
    if rc:
        n = C.c_size_t()
        rtc.hiprtcGetProgramLogSize(prog, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        rtc.hiprtcGetProgramLog(prog, buf)
        log = buf.value
    rtc.hiprtcDestroyProgram(C.byref(prog))
    return rc, log


@pytest.mark.parametrize("dtype,regs", [("c128", 16), ("c64", 32)])
def test_source_structure(dtype, regs):
    prog = qlx.layered(14, 4, 7)
    order, sweep_of, info = qlx.plan_windowed(prog, dtype, qlx.exec_opts("windowed"))
    assert info["sweeps"] >= 1
    for s in range(info["sweeps"]):
        src = qlx.sweep_source(prog, s, dtype)
        assert re.search(rf"qlx_sweep_spec_[0-9a-f]{{6}}_{s}\(", src)
        assert len(re.findall(r"= amp\[lo \| 0x[0-9a-f]+ull\];", src)) == regs
        # stores: through `so`, or `sf` = so ^ (the pending work-item-controlled X of the registers)
        assert len(re.findall(r"amp\[(?:so \||sf \^) 0x[0-9a-f]+ull\] = r\d+;", src)) == regs
        # gate coefficients are literals; no descriptor arrays are read
        assert "coef[" not in src and "ops[" not in src
        # every remap writes and reads all registers through LDS
        remaps = src.count("__syncthreads();") // 2
        assert len(re.findall(r"ldsb \+ \(s0f? \^ \d+u\)\) = r\d+;", src)) == regs * remaps
        assert len(re.findall(r"= \*reinterpret_cast<const \w+\*>\(ldsb \+ \(s1 [+^] \d+u\)\);", src)) == regs * remaps


def test_cx_and_swap_are_register_renaming():
    c = qlx.Circuit(13).cx(0, 1).swap(1, 2).cx(2, 0)
    prog = c.release()
    src = qlx.sweep_source(prog, 0, "c128")
    assert "const double2 a0" not in src          # no arithmetic at all
    stores = re.findall(r"amp\[so \| (0x[0-9a-f]+)ull\] = r(\d+);", src)
    assert sorted(int(v) for _, v in stores) == list(range(16))
    # the renamed registers show up where the chunk leaves the registers (an LDS re-layout or
    # the final store): some register index is written from another register's variable
    writes = [int(v) for v in re.findall(r"ldsb \+ \(s0 \^ \d+u\)\) = r(\d+);", src)] + [int(v) for _, v in stores]
    assert any(v != i % 16 for i, v in enumerate(writes))


@pytest.mark.skipif(not os.path.exists(HIPRTC), reason="hipRTC not installed")
def test_sources_compile_for_gfx950():
    from test_oracle import random_program
    rng = np.random.default_rng(11)
    ins, pool, meta = random_program(13, 40, rng)
    rnd = qlx.Program(13, ins, pool if pool is not None else np.zeros(0, complex),
                      meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    fused = qlx.transpile_gpu(qlx.layered(13, 3, 5), max_width=2, rotations=True)   # U4 ops
    for prog, dtype in [(rnd, "c128"), (fused, "c64"), (qlx.hea(13, 2), "c128")]:
        _, _, info = qlx.plan_windowed(prog, dtype, qlx.exec_opts("windowed"))
        for s in range(min(2, info["sweeps"])):
            rc, log = compile_ok(qlx.sweep_source(prog, s, dtype))
            assert rc == 0, log[-2000:]


def test_cx_controlled_by_a_work_item_bit_costs_no_instructions():
    """A CX whose control is outside the registers (a work-item bit of the phase) is recorded as
    a pending X on its target register bit and taken up by the next LDS re-layout's or the final
    store's addresses: no arithmetic, no register swaps; the kernel compiles."""
    n = 12
    c = qlx.Circuit(n)
    for q in range(n - 1):
        c.cx(q, q + 1)                  # a chain: no 4-qubit register set holds every control
    c.rx(7, 0.3).cx(6, 7).rx(7, 1.1)    # RX on a target with a pending X: applied as it is
    prog = c.release()
    # rotations kept as written (merged, the two RX would be one)
    o = qlx.exec_opts("windowed", specialize=True, keep_rotations=True)
    order, sweep_of, info = qlx.plan_windowed(prog, "c128", o)
    assert info["sweeps"] == 1 and sorted(order) == list(range(len(prog)))
    src = qlx.sweep_source(prog, 0, "c128", o)
    assert "__builtin_popcount(t &" in src              # a parity of work-item bits
    assert "fc ?" not in src                            # never applied as register swaps
    assert src.count("__builtin_fma(") == 2 * 8 * 4     # the two RX (four FMAs per pair), nothing else
    if os.path.exists(HIPRTC):
        rc, log = compile_ok(src)
        assert rc == 0, log[-2000:]


def sources(prog):
    o = qlx.exec_opts("windowed", specialize=True)
    return [qlx.sweep_source(prog, s, "c128", o) for s in range(qlx.plan_windowed(prog, "c128", o)[2]["sweeps"])]


def chain(n):
    c = qlx.Circuit(n)
    for q in range(n - 1):
        c.cx(q, q + 1)                  # work-item-controlled CX: pending X frames on the targets
    return c


def test_h_turns_a_pending_x_into_a_pending_z():
    """H after a work-item-controlled CX on the same target: H X = Z H, so the frame becomes a
    pending Z (no register swaps), applied as sign flips before the re-layout or store."""
    n = 12
    c = chain(n)
    for q in range(1, n):
        c.h(q)
    srcs = sources(c.release())
    assert not any("fc ?" in s for s in srcs)
    assert any("zs = " in s and "^ zs)" in s for s in srcs)
    for s in srcs:
        if os.path.exists(HIPRTC):
            rc, log = compile_ok(s)
            assert rc == 0, log[-2000:]


def test_rx_under_a_pending_z_flips_its_angle_per_work_item():
    """RX(theta) Z = Z RX(-theta): under a pending Z the RX coefficient t = tan(theta/2) takes a
    per-work-item sign; the u = cot(theta/2) form (|sin| > |cos|) applies the Z first."""
    n = 12
    c = chain(n)
    for q in range(1, n):
        c.h(q)
    c.rx(5, 0.4)                        # t form
    c.rx(6, 3.0)                        # u form
    srcs = sources(c.release())
    assert any(re.search(r"const double tc = \(\(__builtin_popcount\(t & \d+u\) & 1u\) != 0u\) \? -\(", s)
               for s in srcs)
    if os.path.exists(HIPRTC):
        for s in srcs:
            rc, log = compile_ok(s)
            assert rc == 0, log[-2000:]
    assert not any("fc ?" in s for s in srcs)


def test_pending_frame_is_applied_before_a_dense_gate():
    """A dense 1-qubit gate (RY runs as a 2x2 matrix) commutes with neither X nor Z: the pending
    X is applied first, as conditional swaps of the register pairs."""
    n = 12
    c = chain(n)
    for q in range(1, n):
        c.ry(q, 0.3)
    srcs = sources(c.release())
    assert any("fc ? a1 : a0" in s for s in srcs)


def test_h_and_rx_take_their_factor_out():
    """H runs as a butterfly (no multiplications) and RX as two FMAs per amplitude; their
    common factors are applied once, as one multiplication per register before the stores."""
    prog = qlx.Circuit(13).h(0).h(5).rx(3, 0.7).rx(4, 3.0).release()
    src = qlx.sweep_source(prog, 0, "c128")
    body = [l for l in src.splitlines() if "a0 = " in l]
    assert body and all("*" not in l for l in body)        # fma calls and adds only
    assert sum("__builtin_fma(" in l for l in body) == 16   # 2 RX x 8 pairs x 1 line
    assert len(re.findall(r"r\d+\.x \*= k;", src)) == 16  # the factor, once per register
    k = float.fromhex(re.search(r"const double k = ([^;]+);", src).group(1))
    import math
    # h^2, then c for RX(0.7) (|c| >= |s|) and s for RX(3.0) (|s| > |c|); angles are stored as float
    want = (1 / math.sqrt(2)) ** 2 * math.cos(np.float32(0.7) / 2) * math.sin(np.float32(3.0) / 2)
    assert abs(k - want) <= 1e-15


def test_wider_windows_for_specialised_sweeps():
    prog = qlx.layered(20, 6, 3)
    with pytest.raises(qlx.QlxUnsupported):
        qlx.plan_windowed(prog, "c128", qlx.exec_opts("windowed", window_qubits=13, register_qubits=4))
    o = qlx.exec_opts("windowed", window_qubits=13, register_qubits=4, specialize=True)
    _, _, info = qlx.plan_windowed(prog, "c128", o)
    assert info["window_qubits"] == 13 and info["threads_per_group"] == 512
    assert "__shared__ double2 lds[8192]" in qlx.sweep_source(prog, 0, "c128", o)
    with pytest.raises(qlx.QlxUnsupported):       # 256 KiB of LDS
        qlx.plan_windowed(prog, "c128", qlx.exec_opts("windowed", window_qubits=14, register_qubits=5,
                                                      specialize=True))


def _bits(expr):
    """Address bit positions an expression of the generated source sets: `<< k` terms of the
    work-item / chunk parts, hexadecimal offsets of the registers."""
    return [int(k) for k in re.findall(r"<< (\d+)\)", expr)]


def _work_bits(expr):
    """(work-item bit, address bit) pairs of a generated `tl` / `ts` expression."""
    return [(int(b), int(k)) for b, k in re.findall(r"\(t >> (\d+)\) & 1u\) << (\d+)\)", expr)]


@pytest.mark.parametrize("prog_name", ["layered", "hea", "random", "layered24"])
def test_relabelled_sweeps_address_every_amplitude_once(prog_name):
    """Out-of-place sweeps (relabel): in every generated kernel the load address bits (chunk
    index bits, work-item bits, register bits) and the store address bits each form a
    permutation of the state's qubits, so loads and stores are in bounds and bijective; the
    kernel reads `src` and writes `dst`; work-item bits 0-2 sit on address bits 0-2 at the
    load and at the store (128-byte runs), also when windows leave out qubits 0-2 (the 24-qubit
    layered circuit, whose plan takes such windows)."""
    from test_oracle import random_program
    n = 24 if prog_name == "layered24" else 17
    if prog_name == "layered24":
        prog = qlx.layered(n, 20, 8)
        fixed = qlx.exec_opts("windowed", specialize=True, relabel=True, fixed_windows=True)
        assert qlx.plan_windowed(prog, "c128", fixed)[2]["sweeps"] > qlx.plan_windowed(
            prog, "c128", qlx.exec_opts("windowed", specialize=True, relabel=True))[2]["sweeps"]
    elif prog_name == "layered":
        prog = qlx.layered(n, 8, 5)
    elif prog_name == "hea":
        prog = qlx.hea(n, 4)
    else:
        ins, pool, meta = random_program(n, 150, np.random.default_rng(3))
        prog = qlx.Program(n, ins, pool if pool is not None else np.zeros(0, complex),
                           meta if meta is not None else np.zeros(0, qlx.META_DTYPE))
    o = qlx.exec_opts("windowed", specialize=True, relabel=True)
    _, _, info = qlx.plan_windowed(prog, "c128", o)
    assert info["sweeps"] >= 2
    for s in range(info["sweeps"]):
        src = qlx.sweep_source(prog, s, "c128", o)
        assert "const double2* __restrict__ src, double2* __restrict__ dst" in src
        assert "amp[" not in src
        for base, thr, ptr, idx in (("base", "tl", "src", "lo"), ("sbase", "ts", "dst", "so")):
            chunk = _bits(re.search(rf"const unsigned long long {base} = ([^;]+);", src).group(1))
            work = _bits(re.search(rf"const unsigned long long {thr} = ([^;]+);", src).group(1))
            pat = rf"{ptr}\[{idx} \| (0x[0-9a-f]+)ull\]" if idx == "lo" else rf"{ptr}\[(?:so \||sf \^) (0x[0-9a-f]+)ull\]"
            offs = [int(h, 16) for h in re.findall(pat, src)]
            assert len(offs) == 16 and len(set(offs)) == 16
            regs = [b for b in range(n) if any((o_ >> b) & 1 for o_ in offs)]
            if idx == "so":
                # the per-work-item XOR of a pending X moves a store only among the register offsets
                for h in re.findall(r"fo \^= (0x[0-9a-f]+)ull;", src):
                    assert int(h, 16) in offs and int(h, 16) != 0
            # register offsets are all combinations of 4 address bits
            assert len(regs) == 4 and sorted(offs) == sorted(
                sum(1 << regs[k] for k in range(4) if (i >> k) & 1) for i in range(16))
            assert sorted(chunk + work + regs) == list(range(n)), (s, base)
        # work-item bits 0-2 are address bits 0-2 at both ends (128-byte runs)
        for thr in ("tl", "ts"):
            pairs = _work_bits(re.search(rf"const unsigned long long {thr} = ([^;]+);", src).group(1))
            assert sorted(k for b, k in pairs if b < 3) == [0, 1, 2], (s, thr, pairs)


@pytest.mark.skipif(not os.path.exists(HIPRTC), reason="hipRTC not installed")
def test_relabelled_sources_compile_for_gfx950():
    prog = qlx.layered(16, 6, 3)
    o = qlx.exec_opts("windowed", specialize=True, relabel=True)
    for dtype in ("c128", "c64"):
        rc, log = compile_ok(qlx.sweep_source(prog, 0, dtype, o))
        assert rc == 0, log[-2000:]


def _parity(x):
    return bin(x).count("1") & 1


def _relayouts(src):
    """Per LDS re-layout of a generated out-of-place complex128 sweep (two rounds through half a
    chunk of LDS): per round, the byte addresses written and read by every (work-item, register)
    that takes part in it, from the generated expressions."""
    terms = lambda ln: [(int(b), int(c)) for b, c in re.findall(r"\(t >> (\d+)\) & 1u\)\) & (\d+)u\)", ln)]
    rls, cur, block = [], None, None
    for ln in src.splitlines():
        if "const unsigned s0 =" in ln:
            cur = {"s0": terms(ln), "fx": [], "ops": []}
            rls.append(cur)
            continue
        if cur is None:
            continue
        m = re.search(r"popcount\(t & (\d+)u\) & 1u\) != 0u\) fx \^= (\d+)u;", ln)
        if m:
            cur["fx"].append((int(m.group(1)), int(m.group(2))))
        if "const unsigned s1 =" in ln:
            cur["s1"] = terms(ln)
        m = re.search(r"const unsigned rw = \(s0f? >> (\d+)\) & 1u;", ln)
        if m:
            cur["k"] = int(m.group(1))
        m = re.search(r"if \((rw|rr) == (\d)u\) \{", ln)
        if m:
            block = (m.group(1), int(m.group(2)), 0)
        if ln.strip() == "else {":
            block = (block[0], block[1], 1)
        m = re.search(r"ldsb \+ \(a0 \^ (\d+)u\)\) = r\d+;", ln)
        if m:
            cur["ops"].append(("w", block[1], block[2], "^", int(m.group(1))))
        m = re.search(r"= \*reinterpret_cast<const double2\*>\(ldsb \+ \(a1 ([+^]) (\d+)u\)\);", ln)
        if m:
            cur["ops"].append(("r", block[1], block[2], m.group(1), int(m.group(2))))

    def xor_terms(ts, t):
        a = 0
        for b, c in ts:
            if (t >> b) & 1:
                a ^= c
        return a

    for rl in rls:
        k = rl["k"]
        comp = lambda x: ((x >> (k + 1)) << k) | (x & ((1 << k) - 1))
        s0 = []
        for t in range(256):
            a = xor_terms(rl["s0"], t)
            for msk, cc in rl["fx"]:
                if _parity(t & msk):
                    a ^= cc
            s0.append(a)
        s1 = [xor_terms(rl["s1"], t) for t in range(256)]
        rounds = []
        for rnd in (0, 1):
            ins = []   # (kind, {work-item: byte address})
            for kind, r, side, op, c in rl["ops"]:
                if r != rnd:
                    continue
                base = s0 if kind == "w" else s1
                act = {}
                for t in range(256):
                    if (((base[t] >> k) & 1) == rnd) == (side == 0):
                        b = comp(base[t])
                        act[t] = (b + c) if op == "+" else (b ^ c)
                ins.append((kind, act))
            rounds.append(ins)
        yield rounds


def test_relayout_slot_maps():
    """Every LDS re-layout of the 30-qubit headline circuit's specialised sweeps runs in two rounds
    through half a chunk of LDS: in each round the work-items write 2048 distinct 16-byte slots of
    the 32 KiB buffer and read back exactly those, every register taking part in one round. The
    accesses obey the bank rules of 16-byte LDS accesses measured on gfx950 (writes: the active lanes
    of each group of 8 consecutive lanes differ in slot bits 0-2; reads: in each 32-lane half, the
    active lanes of either parity of lane bits 2-4 differ in slot bits 0-3). Register offsets that
    the reads add (rather than XOR) never carry into the work-item part."""
    prog = qlx.layered(30, 20, 1002)
    opts = qlx.exec_opts("windowed", specialize=True, relabel=True)
    _, _, info = qlx.plan_windowed(prog, "c128", opts)
    n = 0
    for s in range(info["sweeps"]):
        for rounds in _relayouts(qlx.sweep_source(prog, s, "c128", opts)):
            n += 1
            nw = nr = 0
            for ins in rounds:
                ws = sorted(a for kind, act in ins if kind == "w" for a in act.values())
                rs = sorted(a for kind, act in ins if kind == "r" for a in act.values())
                assert ws == [16 * i for i in range(2048)] and rs == ws
                nw += len(ws)
                nr += len(rs)
                for kind, act in ins:
                    for wave in range(4):
                        if kind == "w":
                            groups = [range(g, g + 8) for g in range(0, 64, 8)]
                            bits = 7
                        else:
                            groups = [[L for L in range(h, h + 32) if _parity(L & 28) == side]
                                      for h in (0, 32) for side in (0, 1)]
                            bits = 15
                        for grp in groups:
                            sl = [(act[64 * wave + L] >> 4) & bits for L in grp if 64 * wave + L in act]
                            assert len(set(sl)) == len(sl)
            assert nw == nr == 4096
    assert n >= 40
