"""Lexer and recursive-descent parser for the synthesizable Verilog subset.

Supported: one or more modules (the top is flattened by the elaborator when it
instantiates others), ANSI and non-ANSI port lists, ``parameter``/``localparam``,
``wire``/``reg``/``logic``/``integer`` declarations with packed ranges and one
unpacked dimension (memories), ``assign``, ``always @(posedge clk)``,
``always @*``/``always_comb``/``always_ff``, ``initial``, ``begin/end``,
``if/else``, ``case/casez/casex``, blocking and non-blocking assignments,
constant-bound ``for`` loops, and the full Verilog expression grammar without
real numbers, strings, or ``**`` on non-constants.

The node classes are plain tuples-with-names so the elaborator can pattern-match
them cheaply; every node carries the source line for diagnostics.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import List, Optional, Tuple


class VerilogError(Exception):
    pass


# ----------------------------------------------------------------------------
# Lexer
# ----------------------------------------------------------------------------

_TOKEN_RE = re.compile(
    r"""
    (?P<ws>[ \t\r\f\v]+)
  | (?P<nl>\n)
  | (?P<lcomment>//[^\n]*)
  | (?P<bcomment>/\*.*?\*/)
  | (?P<attr>\(\*.*?\*\))
  | (?P<directive>`[A-Za-z_][A-Za-z0-9_]*[^\n]*)
  | (?P<number>(?:[0-9][0-9_]*)?\s*'[sS]?[bBoOdDhH]\s*[0-9a-fA-FxXzZ?_]+|[0-9][0-9_]*)
  | (?P<sysid>\$[A-Za-z_][A-Za-z0-9_$]*)
  | (?P<id>[A-Za-z_][A-Za-z0-9_$]*|\\\S+)
  | (?P<op><<<=|>>>=|<<<|>>>|===|!==|<<=|>>=|\+:|-:|==|!=|<=|>=|&&|\|\||<<|>>|~&|~\||~\^|\^~|\*\*|->|[-+*/%<>=!~&|^?:;,.#@(){}\[\]])
  | (?P<str>"(?:[^"\\]|\\.)*")
    """,
    re.VERBOSE | re.DOTALL,
)

KEYWORDS = {
    "module", "endmodule", "input", "output", "inout", "wire", "reg", "logic", "integer",
    "signed", "unsigned", "parameter", "localparam", "assign", "always", "always_comb",
    "always_ff", "always_latch", "initial", "begin", "end", "if", "else", "case", "casez",
    "casex", "endcase", "default", "posedge", "negedge", "or", "for", "generate",
    "endgenerate", "genvar", "function", "endfunction", "bit",
}


@dataclass
class Tok:
    kind: str  # 'id', 'kw', 'num', 'op', 'sysid', 'str', 'eof'
    text: str
    line: int


def lex(src: str) -> List[Tok]:
    toks: List[Tok] = []
    pos = 0
    line = 1
    n = len(src)
    m = _TOKEN_RE.match
    while pos < n:
        mo = m(src, pos)
        if mo is None:
            raise VerilogError(f"line {line}: unexpected character {src[pos]!r}")
        kind = mo.lastgroup
        text = mo.group(kind)
        pos = mo.end()
        if kind == "nl":
            line += 1
            continue
        if kind in ("ws", "lcomment", "attr"):
            continue
        if kind == "bcomment":
            line += text.count("\n")
            continue
        if kind == "directive":
            if text.startswith("`timescale") or text.startswith("`default_nettype"):
                continue
            raise VerilogError(f"line {line}: compiler directive not supported: {text}")
        if kind == "id":
            toks.append(Tok("kw" if text in KEYWORDS else "id", text, line))
        elif kind == "number":
            toks.append(Tok("num", re.sub(r"\s+", "", text), line))
        else:
            toks.append(Tok(kind, text, line))
    toks.append(Tok("eof", "", line))
    return toks


# ----------------------------------------------------------------------------
# AST
# ----------------------------------------------------------------------------

@dataclass
class Node:
    line: int


@dataclass
class ENum(Node):
    value: int          # value with x/z bits as 0 (two-state)
    width: Optional[int]  # None = unsized (32-bit context)
    signed: bool
    zmask: int = 0      # bits written as z or ? (casez/casex don't-care)
    xmask: int = 0      # bits written as x (casex don't-care)


@dataclass
class EId(Node):
    name: str


@dataclass
class EIndex(Node):       # base[idx]  (bit select or memory element)
    base: Node
    index: Node


@dataclass
class ERange(Node):       # base[msb:lsb]
    base: Node
    msb: Node
    lsb: Node


@dataclass
class EIndexedRange(Node):  # base[start +: width] / base[start -: width]
    base: Node
    start: Node
    width: Node
    up: bool


@dataclass
class EUnary(Node):
    op: str
    a: Node


@dataclass
class EBinary(Node):
    op: str
    a: Node
    b: Node


@dataclass
class ECond(Node):
    c: Node
    a: Node
    b: Node


@dataclass
class EConcat(Node):
    parts: List[Node]


@dataclass
class ERepl(Node):
    count: Node
    parts: List[Node]


@dataclass
class ECall(Node):        # $signed / $unsigned / $clog2 / user function
    name: str
    args: List[Node]


# statements
@dataclass
class SBlock(Node):
    stmts: List[Node]


@dataclass
class SAssign(Node):
    lhs: Node
    rhs: Node
    nonblocking: bool


@dataclass
class SIf(Node):
    cond: Node
    then: Optional[Node]
    els: Optional[Node]


@dataclass
class SCase(Node):
    kind: str  # case / casez / casex
    expr: Node
    items: List[Tuple[Optional[List[Node]], Optional[Node]]]  # (None = default, stmt)


@dataclass
class SFor(Node):
    var: str
    init: Node
    cond: Node
    step: Node
    body: Node


@dataclass
class SNull(Node):
    pass


# module items
@dataclass
class Decl(Node):
    kind: str                    # 'input', 'output', 'inout', 'wire', 'reg', 'integer'
    net: Optional[str]           # for ports: 'wire'/'reg'/None
    signed: bool
    msb: Optional[Node]
    lsb: Optional[Node]
    name: str
    array: Optional[Tuple[Node, Node]]  # unpacked dimension
    init: Optional[Node]


@dataclass
class Param(Node):
    name: str
    value: Node
    local: bool
    msb: Optional[Node] = None
    lsb: Optional[Node] = None
    signed: bool = False


@dataclass
class ContAssign(Node):
    lhs: Node
    rhs: Node


@dataclass
class Always(Node):
    kind: str                    # 'comb' or 'seq'
    edges: List[Tuple[str, str]]  # [('posedge', 'clk')]
    body: Node


@dataclass
class Initial(Node):
    body: Node


@dataclass
class Instance(Node):
    module: str
    name: str
    params: List[Tuple[Optional[str], Node]]
    conns: List[Tuple[Optional[str], Optional[Node]]]


@dataclass
class Module(Node):
    name: str
    ports: List[str]
    items: List[Node] = field(default_factory=list)


# ----------------------------------------------------------------------------
# Number literals
# ----------------------------------------------------------------------------

_BASES = {"b": 2, "o": 8, "d": 10, "h": 16}


def parse_number(text: str, line: int) -> ENum:
    t = text.replace("_", "")
    if "'" not in t:
        return ENum(line, int(t), None, True)
    size_s, rest = t.split("'", 1)
    signed = False
    if rest[0] in "sS":
        signed = True
        rest = rest[1:]
    base = _BASES[rest[0].lower()]
    digits = rest[1:].lower()
    width = int(size_s) if size_s else None
    if base == 10:
        if any(ch in "xz?" for ch in digits):
            value = 0
        else:
            value = int(digits, 10)
    else:
        bits_per = {2: 1, 8: 3, 16: 4}[base]
        value = 0
        zmask = 0
        xmask = 0
        full = (1 << bits_per) - 1
        for ch in digits:
            value <<= bits_per
            zmask <<= bits_per
            xmask <<= bits_per
            if ch in "z?":
                zmask |= full
            elif ch == "x":
                xmask |= full
            else:
                value |= int(ch, base)
        if width is not None:
            zmask &= (1 << width) - 1
            xmask &= (1 << width) - 1
            value &= (1 << width) - 1
        return ENum(line, value, width, signed, zmask, xmask)
    if width is not None:
        value &= (1 << width) - 1
    return ENum(line, value, width, signed)


# ----------------------------------------------------------------------------
# Parser
# ----------------------------------------------------------------------------

# binary operator precedence (higher binds tighter), IEEE 1364-2005 table 5-4
_BINPREC = {
    "||": 1, "&&": 2, "|": 3, "^": 4, "^~": 4, "~^": 4, "&": 5,
    "==": 6, "!=": 6, "===": 6, "!==": 6,
    "<": 7, "<=": 7, ">": 7, ">=": 7,
    "<<": 8, ">>": 8, "<<<": 8, ">>>": 8,
    "+": 9, "-": 9, "*": 10, "/": 10, "%": 10, "**": 11,
}
_UNARY = {"!", "~", "-", "+", "&", "|", "^", "~&", "~|", "~^", "^~"}


class Parser:
    def __init__(self, toks: List[Tok]):
        self.toks = toks
        self.i = 0

    # -- token helpers --
    def peek(self, k: int = 0) -> Tok:
        return self.toks[self.i + k]

    def next(self) -> Tok:
        t = self.toks[self.i]
        self.i += 1
        return t

    def at(self, text: str) -> bool:
        t = self.toks[self.i]
        return t.text == text and t.kind in ("op", "kw")

    def accept(self, text: str) -> bool:
        if self.at(text):
            self.i += 1
            return True
        return False

    def expect(self, text: str) -> Tok:
        t = self.next()
        if t.text != text or t.kind not in ("op", "kw"):
            raise VerilogError(f"line {t.line}: expected {text!r}, found {t.text!r}")
        return t

    def ident(self) -> str:
        t = self.next()
        if t.kind != "id":
            raise VerilogError(f"line {t.line}: expected identifier, found {t.text!r}")
        return t.text

    def err(self, msg: str):
        raise VerilogError(f"line {self.peek().line}: {msg}")

    # -- top level --
    def parse(self) -> List[Module]:
        mods = []
        while self.peek().kind != "eof":
            if self.at("module"):
                mods.append(self.module())
            else:
                self.err(f"expected 'module', found {self.peek().text!r}")
        return mods

    def module(self) -> Module:
        line = self.expect("module").line
        name = self.ident()
        mod = Module(line, name, [])
        if self.accept("#"):
            self.expect("(")
            while not self.at(")"):
                self.accept("parameter")
                mod.items.extend(self.param_list(local=False, terminators=(",", ")")))
                if not self.accept(","):
                    break
            self.expect(")")
        if self.accept("("):
            if not self.at(")"):
                self.port_list(mod)
            self.expect(")")
        self.expect(";")
        while not self.at("endmodule"):
            self.module_item(mod.items)
        self.expect("endmodule")
        return mod

    def port_list(self, mod: Module):
        # ANSI (with direction) or non-ANSI (names only)
        last: Optional[Tuple[str, Optional[str], bool, Optional[Node], Optional[Node]]] = None
        while True:
            t = self.peek()
            if t.text in ("input", "output", "inout"):
                kind = self.next().text
                net = None
                if self.peek().text in ("wire", "reg", "logic"):
                    net = self.next().text
                signed = self.accept("signed")
                msb = lsb = None
                if self.at("["):
                    msb, lsb = self.range_()
                last = (kind, net, signed, msb, lsb)
                pname = self.ident()
                mod.ports.append(pname)
                mod.items.append(Decl(t.line, kind, net, signed, msb, lsb, pname, None, None))
            elif t.kind == "id":
                pname = self.ident()
                mod.ports.append(pname)
                if last is not None:
                    kind, net, signed, msb, lsb = last
                    mod.items.append(Decl(t.line, kind, net, signed, msb, lsb, pname, None, None))
            else:
                self.err("bad port list")
            if not self.accept(","):
                break

    def range_(self) -> Tuple[Node, Node]:
        self.expect("[")
        msb = self.expr()
        self.expect(":")
        lsb = self.expr()
        self.expect("]")
        return msb, lsb

    def param_list(self, local: bool, terminators=(";",)) -> List[Param]:
        out = []
        if self.peek().text in ("integer",):
            self.next()
        signed = self.accept("signed")
        msb = lsb = None
        if self.at("["):
            msb, lsb = self.range_()
        while True:
            line = self.peek().line
            name = self.ident()
            self.expect("=")
            val = self.expr()
            out.append(Param(line, name, val, local, msb, lsb, signed))
            if self.at(",") and self.peek(1).kind == "id" and self.peek(2).text == "=":
                self.next()
                continue
            break
        return out

    def module_item(self, items: List[Node]):
        t = self.peek()
        txt = t.text
        if txt in ("input", "output", "inout", "wire", "reg", "logic", "integer", "bit"):
            self.decl(items)
        elif txt in ("parameter", "localparam"):
            self.next()
            items.extend(self.param_list(local=(txt == "localparam")))
            self.expect(";")
        elif txt == "assign":
            self.next()
            while True:
                line = self.peek().line
                lhs = self.lvalue()
                self.expect("=")
                rhs = self.expr()
                items.append(ContAssign(line, lhs, rhs))
                if not self.accept(","):
                    break
            self.expect(";")
        elif txt in ("always", "always_ff", "always_comb", "always_latch"):
            self.next()
            edges: List[Tuple[str, str]] = []
            kind = "comb"
            if txt in ("always", "always_ff") and self.accept("@"):
                if self.accept("*"):
                    pass
                else:
                    self.expect("(")
                    if self.accept("*"):
                        self.expect(")")
                    else:
                        while True:
                            if self.peek().text in ("posedge", "negedge"):
                                e = self.next().text
                                edges.append((e, self.ident()))
                            else:
                                self.expr()  # level-sensitive list: treated as @*
                            if not (self.accept("or") or self.accept(",")):
                                break
                        self.expect(")")
            if edges:
                kind = "seq"
            body = self.stmt()
            items.append(Always(t.line, kind, edges, body))
        elif txt == "initial":
            self.next()
            items.append(Initial(t.line, self.stmt()))
        elif txt in ("generate", "endgenerate"):
            self.next()
        elif txt == "genvar":
            self.next()
            while True:
                self.ident()
                if not self.accept(","):
                    break
            self.expect(";")
        elif t.kind == "id" and (self.peek(1).kind == "id" or self.peek(1).text == "#"):
            self.instance(items)
        elif txt == ";":
            self.next()
        else:
            self.err(f"unsupported module item starting with {txt!r}")

    def decl(self, items: List[Node]):
        t = self.next()
        kind = t.text
        net = None
        if kind in ("input", "output", "inout"):
            if self.peek().text in ("wire", "reg", "logic"):
                net = self.next().text
        if kind in ("logic", "bit"):
            kind = "reg"
        signed = self.accept("signed")
        if self.accept("unsigned"):
            signed = False
        msb = lsb = None
        if kind == "integer":
            signed = True
            msb = ENum(t.line, 31, None, True)
            lsb = ENum(t.line, 0, None, True)
        elif self.at("["):
            msb, lsb = self.range_()
        while True:
            line = self.peek().line
            name = self.ident()
            arr = None
            if self.at("["):
                arr = self.range_()
            init = None
            if self.accept("="):
                init = self.expr()
            items.append(Decl(line, kind, net, signed, msb, lsb, name, arr, init))
            if not self.accept(","):
                break
        self.expect(";")

    def instance(self, items: List[Node]):
        line = self.peek().line
        mname = self.ident()
        params: List[Tuple[Optional[str], Node]] = []
        if self.accept("#"):
            self.expect("(")
            while not self.at(")"):
                if self.accept("."):
                    pn = self.ident()
                    self.expect("(")
                    params.append((pn, self.expr()))
                    self.expect(")")
                else:
                    params.append((None, self.expr()))
                if not self.accept(","):
                    break
            self.expect(")")
        iname = self.ident()
        self.expect("(")
        conns: List[Tuple[Optional[str], Optional[Node]]] = []
        while not self.at(")"):
            if self.accept("."):
                pn = self.ident()
                self.expect("(")
                e = None if self.at(")") else self.expr()
                self.expect(")")
                conns.append((pn, e))
            else:
                conns.append((None, self.expr()))
            if not self.accept(","):
                break
        self.expect(")")
        self.expect(";")
        items.append(Instance(line, mname, iname, params, conns))

    # -- statements --
    def stmt(self) -> Node:
        t = self.peek()
        txt = t.text
        if txt == "begin":
            self.next()
            if self.accept(":"):
                self.ident()
            stmts = []
            while not self.at("end"):
                stmts.append(self.stmt())
            self.expect("end")
            if self.at(":"):
                self.next()
                self.ident()
            return SBlock(t.line, stmts)
        if txt == "if":
            self.next()
            self.expect("(")
            c = self.expr()
            self.expect(")")
            th = self.stmt()
            el = None
            if self.accept("else"):
                el = self.stmt()
            return SIf(t.line, c, th, el)
        if txt in ("case", "casez", "casex"):
            self.next()
            self.expect("(")
            e = self.expr()
            self.expect(")")
            items = []
            while not self.at("endcase"):
                if self.accept("default"):
                    self.accept(":")
                    items.append((None, self.stmt()))
                    continue
                labels = [self.expr()]
                while self.accept(","):
                    labels.append(self.expr())
                self.expect(":")
                items.append((labels, self.stmt()))
            self.expect("endcase")
            return SCase(t.line, txt, e, items)
        if txt == "for":
            self.next()
            self.expect("(")
            if self.peek().text in ("integer", "int", "genvar"):
                self.next()
            var = self.ident()
            self.expect("=")
            init = self.expr()
            self.expect(";")
            cond = self.expr()
            self.expect(";")
            v2 = self.ident()
            if v2 != var:
                self.err("for-loop step must assign the loop variable")
            self.expect("=")
            step = self.expr()
            self.expect(")")
            body = self.stmt()
            return SFor(t.line, var, init, cond, step, body)
        if txt == ";":
            self.next()
            return SNull(t.line)
        if t.kind == "sysid":
            # $display/$finish etc: ignored on the simulation path
            self.next()
            if self.accept("("):
                depth = 1
                while depth:
                    tt = self.next()
                    if tt.text == "(":
                        depth += 1
                    elif tt.text == ")":
                        depth -= 1
            self.expect(";")
            return SNull(t.line)
        lhs = self.lvalue()
        if self.accept("<="):
            rhs = self.expr()
            self.expect(";")
            return SAssign(t.line, lhs, rhs, True)
        self.expect("=")
        rhs = self.expr()
        self.expect(";")
        return SAssign(t.line, lhs, rhs, False)

    def lvalue(self) -> Node:
        if self.at("{"):
            line = self.next().line
            parts = [self.lvalue()]
            while self.accept(","):
                parts.append(self.lvalue())
            self.expect("}")
            return EConcat(line, parts)
        line = self.peek().line
        e: Node = EId(line, self.ident())
        return self.selects(e)

    def selects(self, e: Node) -> Node:
        while self.at("["):
            line = self.next().line
            a = self.expr()
            if self.accept(":"):
                b = self.expr()
                self.expect("]")
                e = ERange(line, e, a, b)
            elif self.at("+:") or self.at("-:"):
                up = self.next().text == "+:"
                w = self.expr()
                self.expect("]")
                e = EIndexedRange(line, e, a, w, up)
            else:
                self.expect("]")
                e = EIndex(line, e, a)
        return e

    # -- expressions --
    def expr(self) -> Node:
        c = self.binexpr(1)
        if self.at("?"):
            line = self.next().line
            a = self.expr()
            self.expect(":")
            b = self.expr()
            return ECond(line, c, a, b)
        return c

    def binexpr(self, minprec: int) -> Node:
        lhs = self.unary()
        while True:
            t = self.peek()
            prec = _BINPREC.get(t.text) if t.kind == "op" else None
            if prec is None or prec < minprec:
                return lhs
            self.next()
            rhs = self.binexpr(prec + 1)
            op = "~^" if t.text == "^~" else t.text
            lhs = EBinary(t.line, op, lhs, rhs)

    def unary(self) -> Node:
        t = self.peek()
        if t.kind == "op" and t.text in _UNARY:
            self.next()
            op = "~^" if t.text == "^~" else t.text
            return EUnary(t.line, op, self.unary())
        return self.primary()

    def primary(self) -> Node:
        t = self.next()
        if t.kind == "num":
            return parse_number(t.text, t.line)
        if t.kind == "id":
            if self.at("("):
                # user function call
                self.next()
                args = []
                while not self.at(")"):
                    args.append(self.expr())
                    if not self.accept(","):
                        break
                self.expect(")")
                return ECall(t.line, t.text, args)
            return self.selects(EId(t.line, t.text))
        if t.kind == "sysid":
            args = []
            if self.accept("("):
                while not self.at(")"):
                    args.append(self.expr())
                    if not self.accept(","):
                        break
                self.expect(")")
            return ECall(t.line, t.text, args)
        if t.text == "(":
            e = self.expr()
            self.expect(")")
            return e
        if t.text == "{":
            first = self.expr()
            if self.at("{"):
                # replication {n{a, b}}
                self.next()
                parts = [self.expr()]
                while self.accept(","):
                    parts.append(self.expr())
                self.expect("}")
                self.expect("}")
                return self.selects(ERepl(t.line, first, parts))
            parts = [first]
            while self.accept(","):
                parts.append(self.expr())
            self.expect("}")
            return self.selects(EConcat(t.line, parts))
        raise VerilogError(f"line {t.line}: unexpected token {t.text!r} in expression")


def parse_verilog(src: str) -> List[Module]:
    return Parser(lex(src)).parse()
