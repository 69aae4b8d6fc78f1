"""Generated C++ model class ``V<top>`` over the runtime C ABI.

The reference emits a ``Vtop`` class whose public members are port
descriptors (``RF_IN8(name, msb, lsb)`` etc., src/V3EmitC.cpp:2165-2193,
:2272-2295) and whose ``eval()`` runs the shared ``RTLflow`` graph
(src/V3EmitC.cpp:3663-3735); the testbench writes a stimulus through
``rtlflow.get(top->port, lane)``. ``emit_vtop_header`` writes the same surface
for a design built here: ``V<top>::<port>`` descriptors, ``eval()`` /
``eval_step()`` / ``eval_end_step()`` / ``final()``, per-lane ``set``/``get``
and bulk port transfers, plus the on-device cycle protocol and the batched
work queue. Like the reference model, the class throws std::runtime_error on
failure; the C ABI underneath never throws.

``emit_testbench`` writes a host-driven testbench that uses only that class:
it generates every lane's inputs on the host with the stimulus generator,
drives the clock with two ``eval()`` calls per cycle and folds the outputs
into per-lane FNV-1a-64 hashes -- the way a reference user drives a model --
or runs the same protocol on the device (``begin`` + ``run_cycles``).
"""
from __future__ import annotations

from typing import List

CLS_BYTES = {0: 1, 1: 2, 2: 4, 3: 8}
CLS_CTYPE = {0: "uint8_t", 1: "uint16_t", 2: "uint32_t", 3: "uint64_t"}


def _elem_bytes(p: dict) -> int:
    return CLS_BYTES[p["cls"]] * (p["words"] if p["width"] > 64 else 1)


def emit_vtop_header(manifest: dict) -> str:
    name = manifest["name"]
    cls = f"V{name}"
    ports = manifest["ports"]
    o: List[str] = []
    o.append(f"// {cls}.h -- model class of design '{name}' for batched stimuli on gfx950.")
    o.append("// Generated; link with libV" + name + ".so and librf_verilated.so.")
    o.append("#pragma once")
    o.append("#include <cstdint>")
    o.append("#include <cstring>")
    o.append("#include <stdexcept>")
    o.append("#include <string>")
    o.append("#include <vector>")
    o.append('#include "rf_verilated.h"')
    o.append("")
    o.append(f'extern "C" const rf_design* rf_design_get_{name}(void);')
    o.append("")
    o.append("#ifndef RF_PORT_DESC_DEFINED")
    o.append("#define RF_PORT_DESC_DEFINED")
    o.append("// A port: index into the design's port table, width in bits, storage")
    o.append("// words per lane (u32 words for ports wider than 64 bits), bytes per lane.")
    o.append("struct rf_port_desc {")
    o.append("    uint32_t index, width, words, bytes;")
    o.append("    uint32_t dir;  // 0 data input, 1 output, 2 clock")
    o.append("};")
    o.append("#endif")
    o.append("")
    o.append(f"class {cls} {{")
    o.append("  public:")
    o.append("    // ports")
    for i, p in enumerate(ports):
        o.append(f"    static constexpr rf_port_desc {p['name']}{{{i}u, {p['width']}u, {p['words']}u, "
                 f"{_elem_bytes(p)}u, {p['dir']}u}};")
    o.append(f"    static constexpr uint32_t n_ports = {len(ports)}u;")
    o.append("")
    o.append(f"    explicit {cls}(uint32_t n_lanes, const rf_config* cfg = nullptr) {{")
    o.append(f"        check(rf_create(rf_design_get_{name}(), n_lanes, cfg, &m_));")
    o.append("    }")
    o.append(f"    ~{cls}() {{ if (m_) rf_destroy(m_); }}")
    o.append(f"    {cls}(const {cls}&) = delete;")
    o.append(f"    {cls}& operator=(const {cls}&) = delete;")
    o.append("")
    o.append("    uint32_t lanes() const { return rf_lanes(m_); }")
    o.append("    rf_model* handle() { return m_; }")
    o.append("    static const char* name() { return \"" + name + "\"; }")
    o.append("")
    o.append("    // one lane's value of a port up to 64 bits wide")
    o.append("    void set(const rf_port_desc& p, uint32_t lane, uint64_t v) {")
    o.append("        if (p.width > 64) throw std::invalid_argument(\"set: wide port, use set_words\");")
    o.append("        check(rf_set_port(m_, p.index, lane, 1, &v));  // little-endian low bytes")
    o.append("    }")
    o.append("    uint64_t get(const rf_port_desc& p, uint32_t lane) {")
    o.append("        if (p.width > 64) throw std::invalid_argument(\"get: wide port, use get_words\");")
    o.append("        uint64_t v = 0;")
    o.append("        check(rf_get_port(m_, p.index, lane, 1, &v));")
    o.append("        return v;")
    o.append("    }")
    o.append("    // one lane's value of a wide port as p.words u32 words, word 0 least significant")
    o.append("    void set_words(const rf_port_desc& p, uint32_t lane, const uint32_t* w) {")
    o.append("        check(rf_set_port(m_, p.index, lane, 1, w));")
    o.append("    }")
    o.append("    void get_words(const rf_port_desc& p, uint32_t lane, uint32_t* w) {")
    o.append("        check(rf_get_port(m_, p.index, lane, 1, w));")
    o.append("    }")
    o.append("    // lanes [lane0, lane0+count): p.bytes per lane, lane-major")
    o.append("    void set_bulk(const rf_port_desc& p, uint32_t lane0, uint32_t count, const void* src) {")
    o.append("        check(rf_set_port(m_, p.index, lane0, count, src));")
    o.append("    }")
    o.append("    void get_bulk(const rf_port_desc& p, uint32_t lane0, uint32_t count, void* dst) {")
    o.append("        check(rf_get_port(m_, p.index, lane0, count, dst));")
    o.append("    }")
    o.append("    void set_done(uint32_t lane0, uint32_t count, const uint8_t* done) {")
    o.append("        check(rf_set_done(m_, lane0, count, done));")
    o.append("    }")
    o.append("")
    o.append("    // evaluate every active lane until settled")
    o.append("    void eval() { check(rf_eval(m_)); }")
    o.append("    void eval_step() { eval(); }")
    o.append("    void eval_end_step() {}")
    o.append("    void final() { check(rf_sync(m_)); }")
    o.append("")
    o.append("    // on-device clock-cycle protocol and output hashing")
    o.append("    void begin(uint64_t seed, uint32_t stim0 = 0) { check(rf_begin(m_, seed, stim0)); }")
    o.append("    void run_cycles(uint32_t n, uint32_t mode = RF_MODE_FUSED) {")
    o.append("        check(rf_run_cycles(m_, n, mode));")
    o.append("    }")
    o.append("    // user stimulus / output capture: device buffers [cycle][word][stimulus]")
    o.append("    void set_input_stream(const uint64_t* dev, uint32_t nstim, uint32_t ncycles,")
    o.append("                          uint64_t stim0 = 0, uint32_t cycle0 = 0) {")
    o.append("        check(rf_set_input_stream(m_, dev, stim0, nstim, cycle0, ncycles));")
    o.append("    }")
    o.append("    void set_output_capture(uint64_t* dev, uint32_t nstim, uint32_t ncycles,")
    o.append("                            uint64_t stim0 = 0, uint32_t cycle0 = 0) {")
    o.append("        check(rf_set_output_capture(m_, dev, stim0, nstim, cycle0, ncycles));")
    o.append("    }")
    o.append("    std::vector<uint64_t> hashes() {")
    o.append("        std::vector<uint64_t> h(lanes());")
    o.append("        check(rf_hashes(m_, 0, (uint32_t)h.size(), h.data()));")
    o.append("        return h;")
    o.append("    }")
    o.append("    // batched form: stimuli with per-stimulus cycle budgets through a work queue")
    o.append("    std::vector<uint64_t> submit_queue(uint64_t seed, const std::vector<uint32_t>& budgets) {")
    o.append("        std::vector<uint64_t> h(budgets.size());")
    o.append("        check(rf_submit_queue(m_, seed, budgets.data(), (uint32_t)budgets.size(), nullptr,")
    o.append("                              h.data()));")
    o.append("        return h;")
    o.append("    }")
    o.append("")
    o.append("  private:")
    o.append("    static void check(int rc) {")
    o.append("        if (rc) throw std::runtime_error(std::string(\"" + cls + ": \") + rf_last_error());")
    o.append("    }")
    o.append("    rf_model* m_ = nullptr;")
    o.append("};")
    return "\n".join(o) + "\n"


def emit_testbench(manifest: dict) -> str:
    """``V<top>_tb``: usage ``V<top>_tb LANES CYCLES SEED host|device``;
    prints one hash per lane (hex)."""
    name = manifest["name"]
    cls = f"V{name}"
    ports = manifest["ports"]
    o: List[str] = []
    o.append(f"// {cls}_tb.cpp -- host-driven testbench of design '{name}' (generated)")
    o.append(f'#include "{cls}.h"')
    o.append('#include "rf_stim.h"')
    o.append("#include <cinttypes>")
    o.append("#include <cstdio>")
    o.append("#include <cstdlib>")
    o.append("")
    o.append("int main(int argc, char** argv) {")
    o.append("    if (argc != 5) { std::fprintf(stderr, \"usage: %s LANES CYCLES SEED host|device\\n\", argv[0]); return 2; }")
    o.append("    const uint32_t lanes = (uint32_t)std::strtoul(argv[1], nullptr, 0);")
    o.append("    const uint32_t cycles = (uint32_t)std::strtoul(argv[2], nullptr, 0);")
    o.append("    const uint64_t seed = std::strtoull(argv[3], nullptr, 0);")
    o.append("    const bool host = std::string(argv[4]) == \"host\";")
    o.append("    try {")
    o.append(f"        {cls} top(lanes);")
    o.append("        std::vector<uint64_t> h(lanes, RF_FNV_BASIS);")
    o.append("        if (!host) {")
    o.append("            top.begin(seed, 0);")
    o.append("            top.run_cycles(cycles, RF_MODE_FUSED);")
    o.append("            h = top.hashes();")
    o.append("        } else {")
    o.append("            std::vector<uint8_t> buf;")
    o.append("            for (uint32_t c = 0; c < cycles; ++c) {")
    for p in ports:
        if p["dir"] == 0:
            eb = _elem_bytes(p)
            o.append(f"                buf.assign((size_t)lanes * {eb}u, 0);")
            o.append("                for (uint32_t s = 0; s < lanes; ++s) {")
            if p["width"] > 64:
                chunks = (p["width"] + 63) // 64
                o.append(f"                    uint32_t w[{p['words']}] = {{0}};")
                for j in range(chunks):
                    o.append(f"                    {{ const uint64_t r = rf_stim(seed, s, c, "
                             f"{(p['stim_index'] << 3) | j}u);")
                    o.append(f"                      w[{2 * j}] = (uint32_t)r;")
                    if 2 * j + 1 < p["words"]:
                        o.append(f"                      w[{2 * j + 1}] = (uint32_t)(r >> 32);")
                    o.append("                    }")
                rem = p["width"] % 32
                if rem:
                    o.append(f"                    w[{p['words'] - 1}] &= 0x{(1 << rem) - 1:x}u;")
                o.append(f"                    std::memcpy(&buf[(size_t)s * {eb}u], w, {eb});")
            else:
                mask = (1 << p["width"]) - 1
                ct = CLS_CTYPE[p["cls"]]
                o.append(f"                    const {ct} v = ({ct})(rf_stim(seed, s, c, "
                         f"{p['stim_index'] << 3}u) & 0x{mask:x}ull);")
                o.append(f"                    std::memcpy(&buf[(size_t)s * {eb}u], &v, {eb});")
            o.append("                }")
            o.append(f"                top.set_bulk({cls}::{p['name']}, 0, lanes, buf.data());")
    clocks = [p for p in ports if p["dir"] == 2]
    for val in (0, 1):
        for p in clocks:
            eb = _elem_bytes(p)
            o.append(f"                buf.assign((size_t)lanes * {eb}u, 0);")
            if val:
                o.append(f"                for (uint32_t s = 0; s < lanes; ++s) buf[(size_t)s * {eb}u] = 1;")
            o.append(f"                top.set_bulk({cls}::{p['name']}, 0, lanes, buf.data());")
        o.append("                top.eval();")
    for p in ports:
        if p["dir"] == 1:
            eb = _elem_bytes(p)
            o.append(f"                buf.assign((size_t)lanes * {eb}u, 0);")
            o.append(f"                top.get_bulk({cls}::{p['name']}, 0, lanes, buf.data());")
            o.append("                for (uint32_t s = 0; s < lanes; ++s)")
            o.append(f"                    for (uint32_t b = 0; b < {eb}u; ++b)")
            o.append(f"                        h[s] = rf_fnv_bytes(h[s], buf[(size_t)s * {eb}u + b], 1);")
    o.append("            }")
    o.append("        }")
    o.append("        top.final();")
    o.append("        for (uint32_t s = 0; s < lanes; ++s) std::printf(\"%016\" PRIx64 \"\\n\", h[s]);")
    o.append("    } catch (const std::exception& e) {")
    o.append("        std::fprintf(stderr, \"%s\\n\", e.what());")
    o.append("        return 1;")
    o.append("    }")
    o.append("    return 0;")
    o.append("}")
    return "\n".join(o) + "\n"
