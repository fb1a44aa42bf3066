"""Reader for the machine-written Lua data files of the gas and kinetics tools.

The reference stores gas models and reaction mechanisms as Lua scripts made only of
assignments of literal tables (``db.N2.M = 0.0280134``, ``reaction[1] = {...}``) and of
table-call statements (``Reaction{...}`` in mechanism inputs).  This reader evaluates
that subset -- numbers, strings, booleans, nil, table constructors with positional,
``name=``, ``['str']=`` and ``[n]=`` fields, dotted/indexed assignment targets and
``Name{...}`` calls -- without running a Lua interpreter.  Used once, to write the JSON
data files; nothing at run time reads Lua.
"""
from __future__ import annotations

import re

_TOKEN = re.compile(r"""
    (?P<ws>\s+) |
    (?P<comment>--\[\[.*?\]\]|--[^\n]*) |
    (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?) |
    (?P<str>'(?:[^'\\]|\\.)*'|"(?:[^"\\]|\\.)*") |
    (?P<name>[A-Za-z_][A-Za-z_0-9]*) |
    (?P<op>==|[{}\[\]=,;.()\-+*/])
""", re.S | re.X)


def _tokens(text):
    out = []
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError(f"cannot tokenise at {text[pos:pos + 40]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind in ("ws", "comment"):
            continue
        out.append((kind, m.group(kind)))
    out.append(("eof", ""))
    return out


class _Parser:
    def __init__(self, text):
        self.t = _tokens(text)
        self.i = 0
        self.env = {}
        self.calls = []

    def peek(self, k=0):
        return self.t[self.i + k]

    def take(self, val=None):
        tok = self.t[self.i]
        if val is not None and tok[1] != val:
            raise ValueError(f"expected {val!r}, got {tok!r}")
        self.i += 1
        return tok

    def expr(self):
        kind, val = self.peek()
        if val == "-":
            self.take()
            return -self.expr()
        if kind == "num":
            self.take()
            v = float(val)
            return int(v) if re.fullmatch(r"\d+", val) else v
        if kind == "str":
            self.take()
            return val[1:-1]
        if val == "{":
            return self.table()
        if kind == "name":
            self.take()
            if val in ("true", "false"):
                return val == "true"
            if val == "nil":
                return None
            obj = self.env.get(val)
            while self.peek()[1] in (".", "["):
                if self.take()[1] == ".":
                    obj = obj[self.take()[1]]
                else:
                    key = self.expr()
                    self.take("]")
                    obj = obj[key]
            return obj
        raise ValueError(f"unexpected token {val!r}")

    def table(self):
        self.take("{")
        pos, named = [], {}
        while self.peek()[1] != "}":
            kind, val = self.peek()
            if val == "[":
                self.take()
                key = self.expr()
                self.take("]")
                self.take("=")
                named[key] = self.expr()
            elif kind == "name" and self.peek(1)[1] == "=":
                self.take()
                self.take("=")
                named[val] = self.expr()
            else:
                pos.append(self.expr())
            if self.peek()[1] in (",", ";"):
                self.take()
        self.take("}")
        if named and not pos:
            return named
        if pos and not named:
            return pos
        d = dict(named)
        for k, v in enumerate(pos, start=1):
            d[k] = v
        return d

    def statement(self):
        kind, val = self.peek()
        if val == "local":
            self.take()
            return self.statement()
        name = self.take()[1]
        if self.peek()[1] == "{":  # Name{...} call
            self.calls.append((name, self.table()))
            return
        if self.peek()[1] == "(":
            raise ValueError("function calls with parentheses are not data")
        path = [name]
        while self.peek()[1] in (".", "["):
            if self.take()[1] == ".":
                path.append(self.take()[1])
            else:
                key = self.expr()
                self.take("]")
                path.append(key)
        self.take("=")
        value = self.expr()
        obj = self.env
        for p in path[:-1]:
            obj = obj[p]
        obj[path[-1]] = value

    def run(self):
        while self.peek()[0] != "eof":
            if self.peek()[1] == ";":
                self.take()
                continue
            self.statement()
        return self.env, self.calls


def parse(text: str):
    """Returns (globals dict, list of (name, table) calls)."""
    return _Parser(text).run()


def parse_file(path: str):
    with open(path) as f:
        return parse(f.read())
