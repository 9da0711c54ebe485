"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function the public headers declare, the ctypes mirrors of the
reference structs match the C layout, argument validation of the drop-in
entry points, and the task runtime's unit bookkeeping
(cellspe-tasklib.cpp:110-412), which runs on host threads.
"""
import ctypes as C
import os
import re
import subprocess
import threading
import time

import numpy as np
import pytest

import fxlib
import tasklib as TL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
HEADERS = sorted(f for f in os.listdir(INC) if f.endswith(".h"))


def declared_functions(header):
    """Function names declared in a header, read through the C preprocessor
    (C mode, so the C++-only helper class is left out)."""
    out = subprocess.run(["gcc", "-E", "-x", "c", "-I", INC, os.path.join(INC, header)],
                         check=True, capture_output=True, text=True).stdout
    # keep only the lines that come from this header (line markers name the file)
    keep, mine = [], False
    for line in out.split("\n"):
        m = re.match(r'^# \d+ "([^"]+)"', line)
        if m:
            mine = os.path.basename(m.group(1)) == header
        elif mine:
            keep.append(line)
    text = re.sub(r"\s+", " ", "\n".join(keep))
    text = re.sub(r"__attribute__\s*\(\((?:[^()]|\([^()]*\))*\)\)", "", text)
    names = set()
    for stmt in text.split(";"):
        stmt = stmt.strip()
        if not stmt or stmt.startswith("typedef") or "{" in stmt or "}" in stmt:
            # skip type definitions; a declaration after a closing brace
            stmt = stmt.split("}")[-1].strip()
            if not stmt or stmt.startswith("typedef"):
                continue
        m = re.match(r"^(?:extern )?(?:const )?[A-Za-z_][\w ]*?[\s\*]+\**([A-Za-z_]\w*)\s*\(", stmt)
        if m:
            names.add(m.group(1))
    return names


def _own_header_functions(header):
    names = declared_functions(header)
    if header != "cellspe_fx.h":
        names -= declared_functions("cellspe_fx.h")
    if header == "celldsp_dsp.h":
        names -= declared_functions("celldsp_tasklib.h")
    return names


@pytest.mark.parametrize("header", HEADERS)
def test_library_exports_every_declared_function(header):
    lib = fxlib.load()
    names = _own_header_functions(header)
    assert names, f"no declarations parsed from {header}"
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_header_parser_sees_known_entry_points():
    assert {"fx_corr_process", "fx_spectrometer_ex", "fx_pcal_final"} <= declared_functions("cellspe_fx.h")
    assert {"celldsp_executeTask", "celldsp_setBehaviour"} <= declared_functions("celldsp_tasklib.h")
    assert {"speDIFX_process", "speUnpack2bit"} <= declared_functions("celldsp_dsp.h")
    assert {"minicorr_run"} <= declared_functions("minicorrelator_fx.h")
    assert {"spectrometer_cb_run", "fx_taskcore_run"} <= declared_functions("spectrometer_fx.h")


# ---------------------------------------------------------------- struct layouts
LAYOUT_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "cellspe_fx.h"
#include "celldsp_tasklib.h"
#include "minicorrelator_fx.h"
#include "spectrometer_fx.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
#define S(T) printf(#T " %zu\n", sizeof(T));
int main(void) {
  S(CellDSPTask_t) P(CellDSPTask_t, completed_flag) P(CellDSPTask_t, tick_count)
  P(CellDSPTask_t, inputvector_ptrs) P(CellDSPTask_t, inputvector_lengths)
  P(CellDSPTask_t, num_output_vectors) P(CellDSPTask_t, outputvector_ptrs)
  P(CellDSPTask_t, outputvector_lengths) P(CellDSPTask_t, subcontext) P(CellDSPTask_t, next_task)
  S(control_block) P(control_block, baselines_out) P(control_block, fmultipliers_src)
  P(control_block, spe_num) P(control_block, ffts_total) P(control_block, ffts_to_integrate)
  P(control_block, spe_sig_listOnPPU)
  S(control_block_t) P(control_block_t, fft_points) P(control_block_t, rawif_src)
  P(control_block_t, rawif_bytecount) P(control_block_t, spectrum_dst)
  S(spectrometer_mbox_t) S(minicorr_stats)
  S(fx_spectrometer_settings) P(fx_spectrometer_settings, samplingfreq)
  P(fx_spectrometer_settings, pcalharmonicshz)
  S(fx_corr_config) P(fx_corr_config, prec) S(fx_phase_model)
  S(fx_swspec_settings) P(fx_swspec_settings, format) P(fx_swspec_settings, use_channel_file)
  P(fx_swspec_settings, spectra_per_chunk) P(fx_swspec_settings, basefilename)
  S(fx_swspec_stats) P(fx_swspec_stats, realtime_factor)
  S(fx_vdif_header) P(fx_vdif_header, nchan) P(fx_vdif_header, second_of_day)
  S(fx_mark5b_header) P(fx_mark5b_header, frac_ns)
  return 0;
}
"""


def test_ctypes_layouts_match_c(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(LAYOUT_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=gnu99", "-I", INC, str(src), "-o", str(exe)], check=True)
    c = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                             check=True).stdout.split("\n") if line)
    c = {k: int(v) for k, v in c.items()}
    py = {
        "CellDSPTask_t": TL.CellDSPTask, "control_block": TL.MiniControlBlock,
        "control_block_t": TL.SpecControlBlock, "fx_spectrometer_settings": TL.SpectrometerSettings,
        "fx_corr_config": fxlib.CorrConfig, "fx_swspec_settings": TL.SwspecSettings,
        "fx_swspec_stats": TL.SwspecStats, "fx_vdif_header": fxlib.VdifHeader,
        "fx_mark5b_header": fxlib.Mark5bHeader,
    }
    for key, val in c.items():
        if "." in key:
            t, f = key.split(".")
            if t in py:
                assert getattr(py[t], f).offset == val, key
    # C pads the 128-byte aligned structs to a multiple of 128; the mirrors
    # carry the same padding so arrays of them have the C stride
    assert c["CellDSPTask_t"] % 128 == 0 and C.sizeof(TL.CellDSPTask) == c["CellDSPTask_t"]
    assert c["control_block"] % 128 == 0 and C.sizeof(TL.MiniControlBlock) == c["control_block"]
    assert C.sizeof(TL.SpecControlBlock) == c["control_block_t"]
    assert C.sizeof(TL.SpectrometerSettings) == c["fx_spectrometer_settings"]
    assert C.sizeof(TL.SpecMbox) == c["spectrometer_mbox_t"]
    assert C.sizeof(TL.MiniStats) == c["minicorr_stats"]
    assert C.sizeof(fxlib.PhaseModel) == c["fx_phase_model"]
    for name in ("fx_swspec_settings", "fx_swspec_stats", "fx_vdif_header", "fx_mark5b_header"):
        assert C.sizeof(py[name]) == c[name], name


# ---------------------------------------------------------------- validation
def test_error_paths_without_device():
    lib = TL.load()
    # control block: terminate flag returns at once, bad sizes are refused
    raw = np.zeros(4096, np.uint8)
    with pytest.raises(fxlib.FxError):
        TL.spectrometer_cb_run(raw, 1000, 1, 4)
    with pytest.raises(fxlib.FxError):
        TL.spectrometer_cb_run(raw, 1024, 3, 4)
    out, mb = TL.spectrometer_cb_run(raw, 1024, 2, 4, terminate=True)
    assert (mb.num_spectra, mb.total_ffts) == (0, 0)
    # TaskCore settings
    s = TL.SpectrometerSettings(fft_points=1024, fft_overlap_factor=2, core_overlapped_ffts=4, window=1,
                                bits_per_sample=2, source_channels=1, use_channel=0, extract_pcal=0,
                                samplingfreq=16e6, pcaloffsethz=10e3, pcalharmonicshz=1e6)
    h = C.c_void_p()
    for field, bad in (("fft_points", 1000), ("bits_per_sample", 8), ("source_channels", 3),
                       ("use_channel", 2), ("window", 9)):
        old = getattr(s, field)
        setattr(s, field, bad)
        assert lib.fx_taskcore_create(0, C.byref(s), C.byref(h)) != 0
        assert TL.last_error()
        setattr(s, field, old)
    # 20 kHz offset at 16 MHz: rotator 800 samples, a multiple of 16 tone bins
    s.extract_pcal, s.source_channels = 1, 2
    assert lib.fx_taskcore_create(0, C.byref(s), C.byref(h)) != 0
    # minicorrelator: station count
    cbs = (TL.MiniControlBlock * 1)()
    assert lib.minicorr_run(cbs, 1, None) == -1
    assert "stations" in TL.last_error()
    # spectrometer application: two inputs need two sink names (both sinks
    # of one name would be opened twice and overwrite each other)
    sw = TL.SwspecSettings()
    sw.core = TL.SpectrometerSettings(fft_points=1024, fft_overlap_factor=1, core_overlapped_ffts=4, window=0,
                                      bits_per_sample=2, source_channels=1, use_channel=0, extract_pcal=0,
                                      samplingfreq=16e6, pcaloffsethz=10e3, pcalharmonicshz=1e6)
    sw.format, sw.spectra_per_chunk = fxlib.FX_FMT_RAW, 1
    sw.use_channel_file[0] = sw.use_channel_file[1] = 1
    for second in (None, b"same"):
        sw.basefilename[0], sw.basefilename[1] = b"same", second
        assert lib.fx_swspectrometer_run(C.byref(sw), b"in1", b"in2", None) == -1   # FX_ERR_ARG
        assert "sink names" in TL.last_error()


def test_product_path_fails_loudly_without_device():
    if fxlib.load().fx_device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(fxlib.FxError):
        fxlib.Context(0)


# ---------------------------------------------------------------- task runtime
@pytest.fixture
def runtime():
    lib = TL.load()
    assert lib.celldsp_start() == 0
    yield lib
    lib.celldsp_setBehaviour(TL.CDSP_BEHAV_DEFAULT)
    while lib.celldsp_clientCount() > 0:
        lib.celldsp_stop()


def test_start_stop_refcount(runtime):
    lib = runtime
    assert lib.celldsp_clientCount() == 1
    assert lib.celldsp_start() == 0
    assert lib.celldsp_clientCount() == 2
    assert lib.celldsp_stop() == 0
    assert lib.celldsp_clientCount() == 1
    assert lib.celldsp_freeCount() == TL.CELLDSP_MAX_SPE_NR


def test_stop_without_start_fails():
    lib = TL.load()
    while lib.celldsp_clientCount() > 0:
        lib.celldsp_stop()
    assert lib.celldsp_stop() == -1
    assert lib.celldsp_getFreeSPE() == -1


def test_reservation_exhaustion_and_release(runtime):
    lib = runtime
    got = [lib.celldsp_getFreeSPE() for _ in range(TL.CELLDSP_MAX_SPE_NR)]
    assert sorted(got) == list(range(TL.CELLDSP_MAX_SPE_NR))
    assert lib.celldsp_freeCount() == 0
    assert lib.celldsp_getFreeSPE() == -1          # non-blocking: none left
    t = TL.CellDSPTaskClass(TL.CELLDSP_TASK_NOP)
    assert lib.celldsp_executeTask(t.getTask()) == 1
    for u in got:
        assert lib.celldsp_setFreeSPE(u) == 0
    assert lib.celldsp_setFreeSPE(got[0]) == 0     # already free: warning only
    assert lib.celldsp_freeCount() == TL.CELLDSP_MAX_SPE_NR
    assert lib.celldsp_setFreeSPE(TL.CELLDSP_MAX_SPE_NR) == -1


def test_start_task_needs_reservation(runtime):
    lib = runtime
    t = TL.CellDSPTaskClass(TL.CELLDSP_TASK_NOP)
    assert lib.celldsp_startTask(3, t.getTask()) == -1
    u = lib.celldsp_getFreeSPE()
    assert lib.celldsp_startTask(u, t.getTask()) == 0
    assert lib.celldsp_waitTask(u) == TL.CELLDSP_TASK_OK
    assert t.completed_flag == TL.CELLDSP_TASK_OK
    lib.celldsp_setFreeSPE(u)


def test_nop_chain_and_unknown_command(runtime):
    lib = runtime
    tasks = [TL.CellDSPTaskClass(TL.CELLDSP_TASK_NOP) for _ in range(4)]
    for a, b in zip(tasks, tasks[1:]):
        a.setNext(b)
    assert lib.celldsp_executeTask(tasks[0].getTask()) == 0
    assert all(t.completed_flag == TL.CELLDSP_TASK_OK for t in tasks)
    bad = TL.CellDSPTaskClass(99)
    assert lib.celldsp_executeTask(bad.getTask()) == 1
    assert bad.completed_flag == TL.CELLDSP_TASK_UNKOWNCOMMAND
    assert lib.speNOP() == 0


def test_continuable_fft_bookkeeping(runtime):
    lib = runtime
    uid = C.c_int(-1)
    assert lib.speFFT_init(3000, C.byref(uid)) == 1      # not a power of two
    assert lib.celldsp_freeCount() == TL.CELLDSP_MAX_SPE_NR
    assert lib.speFFT_init(1024, C.byref(uid)) == 0
    assert 0 <= uid.value < TL.CELLDSP_MAX_SPE_NR
    assert lib.celldsp_freeCount() == TL.CELLDSP_MAX_SPE_NR - 1
    assert lib.speFFT_destroy(C.byref(uid)) == 0
    assert uid.value == -1
    assert lib.celldsp_freeCount() == TL.CELLDSP_MAX_SPE_NR


def test_blocking_get_free_waits(runtime):
    lib = runtime
    held = [lib.celldsp_getFreeSPE() for _ in range(TL.CELLDSP_MAX_SPE_NR)]
    lib.celldsp_setBehaviour(TL.CDSP_BEHAV_BLOCKING)
    assert lib.celldsp_getBehaviour() == TL.CDSP_BEHAV_BLOCKING
    got = []
    th = threading.Thread(target=lambda: got.append(lib.celldsp_getFreeSPE()))
    th.start()
    time.sleep(0.2)
    assert not got                                       # still waiting
    lib.celldsp_setFreeSPE(held[5])
    th.join(timeout=10)
    assert got == [held[5]]
    for u in held:
        lib.celldsp_setFreeSPE(u)


def test_device_commands_report_generic_error_without_device(runtime):
    lib = runtime
    if fxlib.load().fx_device_count() > 0:
        pytest.skip("a HIP device is present")
    out = np.zeros(4, np.float32)
    assert lib.speIOCheck_out(TL._p(out)) == 1
    assert lib.speIOCheck() == 1
    t = TL.CellDSPTaskClass(TL.CELLDSP_TASK_IOTEST)
    t.addOutVec(out)
    assert lib.celldsp_executeTask(t.getTask()) == 1
    assert t.completed_flag == TL.CELLDSP_TASK_GENERICERROR


# ---------------------------------------------------------------- reference declarations
REF = "/root/reference"


def _struct_fields(path, name):
    """Member declarations of `typedef struct ... { ... } name` in a header, as
    normalized (type, declarator) strings: comments, alignment attributes and
    qualifiers that do not change the layout removed."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    m = re.search(r"\}\s*" + name + r"\b[^;]*;", text)
    assert m, (path, name)
    end = m.start()
    depth, i = 0, end
    while True:           # back to the brace that opens this body
        i -= 1
        if text[i] == "}":
            depth += 1
        elif text[i] == "{":
            if depth == 0:
                break
            depth -= 1
    out = []
    for decl in text[i + 1:end].split(";"):
        d = re.sub(r"__attribute__\s*\(\((?:[^()]|\([^()]*\))*\)\)|_QUAD_ALIGN|_CLINE_ALIGN|\bvolatile\b", " ",
                   decl)
        d = " ".join(d.split())
        if d:
            out.append(d.replace(" [", "["))
    return out


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference tree is not present")
@pytest.mark.parametrize("ref_file,name,ours", [
    ("cellspe-tasklib.h", "CellDSPTask_t", "celldsp_tasklib.h"),
    ("minicorrelator/minicorrelator.h", "control_block", "minicorrelator_fx.h"),
    ("spectrometer/src/Cell/cell_spectrometer.h", "control_block_t", "spectrometer_fx.h"),
])
def test_abi_structs_match_reference_declarations(ref_file, name, ours):
    """The drop-in structs declare the reference's members, in order, with the
    same types and array sizes (the reference headers need the Cell SDK, so
    they are compared as declarations rather than compiled)."""
    ref = _struct_fields(os.path.join(REF, ref_file), name)
    got = _struct_fields(os.path.join(INC, ours), name)
    assert got == ref


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "spectrometer/src")),
                    reason="the reference tree is not present")
def test_integration_taskcore_stub_compiles(tmp_path):
    """The TaskCoreHIP class a maintainer adds compiles against
    the reference's own TaskCore.h, Settings.h and Buffer.h and implements
    every pure virtual of TaskCore (it is instantiated through the base)."""
    doc = open(os.path.join(os.path.dirname(INC), "INTEGRATION.md")).read()
    sec = doc[doc.index("## 3. Spectrometer"):]
    code = sec[sec.index("```cpp") + 6: sec.index("```", sec.index("```cpp") + 6)]
    (tmp_path / "TaskCoreHIP.h").write_text("#include <cstring>\n" + code)
    (tmp_path / "main.cpp").write_text('#include "TaskCoreHIP.h"\n'
                                       "int main() { TaskCore* t = new TaskCoreHIP(0); delete t; return 0; }\n")
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-I", str(tmp_path), "-I",
                        os.path.join(REF, "spectrometer/src"), "-I", INC, str(tmp_path / "main.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
