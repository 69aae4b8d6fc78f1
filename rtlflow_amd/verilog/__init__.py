"""Verilog-subset front end: lexer, parser and elaborator producing the design IR.

The reference keeps Verilator's flex/bison front end (src/verilog.l, src/verilog.y,
src/V3Width.cpp); those generators are not available to this build, so this
package parses the synthesizable subset the project's designs use and applies
IEEE 1364-2005 expression sizing rules (section 5.4) itself.
"""
from .parser import parse_verilog  # noqa: F401
from .elab import elaborate  # noqa: F401
