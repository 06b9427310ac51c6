"""Reader for the reference's run configuration files (res/config.lua).

The reference loads `config.lua` into a Lua interpreter and reads globals
from it (src/HydroGPUApp.cpp:164-230): `size`, `xmin`, `xmax`,
`boundaryMethods`, `solverName`, `integratorName`, `slopeLimiter`, `cfl`,
`useFixedDT`, `fixedDT`, `maxFrames`, `useGravity` and the `defs` table
(src/Solver/Solver.cpp:285-307).  The file picks an initial condition by
calling `initConds[name].setup()`, which may itself change those globals
(res/initConds.lua).

This module evaluates the subset of Lua such files use -- literals, table
constructors, global and local assignments (also into table fields),
arithmetic, comparison and boolean operators, `#`, `if`/`elseif`/`else`,
calls of `math.*`, `require` (ignored) and `initConds[...].setup(...)` --
and hands back the globals.  Function definitions are recorded as opaque
values: a config that defines `initState` or `calcSolid` itself is reported,
since this build reproduces the catalogue of initial conditions by name
(initconds.py), whose potential and solid channels come with it.

Lua semantics kept: `--` comments with `--[[ ... ]]` / `--[=[ ... ]=]` block
forms (so `-- [[` toggles a block off, as the reference's config uses it),
1-based table arrays, `nil`/`false` as the only false values, `and`/`or`
returning operands, integer-valued numbers printed as Lua does.
"""
import math
import os
import re

from . import config as _config


class LuaError(ValueError):
    pass


class LuaTable:
    """A Lua table: array part (1-based) plus hash part."""

    def __init__(self, arr=None, hsh=None):
        self.arr = list(arr or [])
        self.hsh = dict(hsh or {})

    def get(self, k):
        if isinstance(k, float) and k.is_integer():
            k = int(k)
        if isinstance(k, int) and 1 <= k <= len(self.arr):
            return self.arr[k - 1]
        return self.hsh.get(k)

    def set(self, k, v):
        if isinstance(k, float) and k.is_integer():
            k = int(k)
        if isinstance(k, int) and 1 <= k <= len(self.arr):
            self.arr[k - 1] = v
            if v is None and k == len(self.arr):
                self.arr.pop()
            return
        if isinstance(k, int) and k == len(self.arr) + 1 and v is not None:
            self.arr.append(v)
            while len(self.arr) + 1 in self.hsh:
                self.arr.append(self.hsh.pop(len(self.arr) + 1))
            return
        if v is None:
            self.hsh.pop(k, None)
        else:
            self.hsh[k] = v

    def length(self):
        return len(self.arr)

    def to_py(self):
        """Arrays become lists, records become dicts (recursively)."""
        if self.hsh and not self.arr:
            return {k: to_py(v) for k, v in self.hsh.items()}
        if self.arr and not self.hsh:
            return [to_py(v) for v in self.arr]
        if not self.arr and not self.hsh:
            return {}
        d = {i + 1: to_py(v) for i, v in enumerate(self.arr)}
        d.update({k: to_py(v) for k, v in self.hsh.items()})
        return d


def to_py(v):
    return v.to_py() if isinstance(v, LuaTable) else v


def from_py(v):
    if isinstance(v, dict):
        return LuaTable(hsh={k: from_py(x) for k, x in v.items()})
    if isinstance(v, (list, tuple)):
        return LuaTable(arr=[from_py(x) for x in v])
    return v


class LuaFunction:
    """A function value: Python callables for builtins; Lua-defined
    functions are kept as opaque (name, source position)."""

    def __init__(self, name, fn=None):
        self.name = name
        self.fn = fn

    def __call__(self, args):
        if self.fn is None:
            raise LuaError("calling the config-defined function %s is not supported" % self.name)
        return self.fn(args)


# ---------------------------------------------------------------- lexer
_TOKEN = re.compile(r"""
    (?P<num>0[xX][0-9a-fA-F]+|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op>\.\.\.|\.\.|==|~=|<=|>=|[-+*/%^#<>=(){}\[\];:,.])
""", re.VERBOSE)

KEYWORDS = {"and", "break", "do", "else", "elseif", "end", "false", "for", "function", "goto", "if", "in",
            "local", "nil", "not", "or", "repeat", "return", "then", "true", "until", "while"}


def _long_bracket(src, i):
    """If src[i:] opens a long bracket [==[, return (level, index after it)."""
    m = re.match(r"\[(=*)\[", src[i:])
    if not m:
        return None
    return len(m.group(1)), i + m.end()


def tokenize(src):
    toks = []
    i, n = 0, len(src)
    while i < n:
        c = src[i]
        if c in " \t\r\n":
            i += 1
            continue
        if src.startswith("--", i):
            lb = _long_bracket(src, i + 2)
            if lb:
                close = "]" + "=" * lb[0] + "]"
                j = src.find(close, lb[1])
                i = n if j < 0 else j + len(close)
            else:
                j = src.find("\n", i)
                i = n if j < 0 else j + 1
            continue
        if c in "\"'":
            j = i + 1
            out = []
            while j < n and src[j] != c:
                if src[j] == "\\" and j + 1 < n:
                    esc = src[j + 1]
                    out.append({"n": "\n", "t": "\t", "\\": "\\", "'": "'", '"': '"'}.get(esc, esc))
                    j += 2
                else:
                    out.append(src[j])
                    j += 1
            if j >= n:
                raise LuaError("unfinished string")
            toks.append(("str", "".join(out)))
            i = j + 1
            continue
        if c == "[":
            lb = _long_bracket(src, i)
            if lb:
                close = "]" + "=" * lb[0] + "]"
                j = src.find(close, lb[1])
                if j < 0:
                    raise LuaError("unfinished long string")
                s = src[lb[1]:j]
                toks.append(("str", s[1:] if s.startswith("\n") else s))
                i = j + len(close)
                continue
        m = _TOKEN.match(src, i)
        if not m:
            raise LuaError("unexpected character %r" % c)
        if m.group("num"):
            t = m.group("num")
            v = float(int(t, 16)) if t.lower().startswith("0x") else float(t)
            toks.append(("num", v))
        elif m.group("name"):
            w = m.group("name")
            toks.append(("kw", w) if w in KEYWORDS else ("name", w))
        else:
            toks.append(("op", m.group("op")))
        i = m.end()
    toks.append(("eof", None))
    return toks


# ---------------------------------------------------------------- parser (to closures)
_BINPRI = {"or": (1, 1), "and": (2, 2), "<": (3, 3), ">": (3, 3), "<=": (3, 3), ">=": (3, 3), "~=": (3, 3),
           "==": (3, 3), "..": (9, 8), "+": (10, 10), "-": (10, 10), "*": (11, 11), "/": (11, 11), "%": (11, 11),
           "^": (14, 13)}
_UNARY_PRI = 12


def truthy(v):
    return v is not None and v is not False


def _num(v):
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        if isinstance(v, str):
            try:
                return float(v)
            except ValueError:
                pass
        raise LuaError("attempt to perform arithmetic on a %s value" % type(v).__name__)
    return float(v)


def _tostr(v):
    if isinstance(v, float) and v.is_integer() and abs(v) < 1e15:
        return str(int(v))
    if isinstance(v, float):
        return "%.14g" % v
    return str(v)


def _binop(op, a, b):
    if op == "+":
        return _num(a) + _num(b)
    if op == "-":
        return _num(a) - _num(b)
    if op == "*":
        return _num(a) * _num(b)
    if op == "/":
        x, y = _num(a), _num(b)
        if y == 0:
            return math.nan if x == 0 or x != x else math.copysign(math.inf, x) * math.copysign(1, y)
        return x / y
    if op == "%":
        x, y = _num(a), _num(b)
        return x - math.floor(x / y) * y
    if op == "^":
        return _num(a) ** _num(b)
    if op == "..":
        return _tostr(a) + _tostr(b)
    if op == "==":
        return a == b and type(a) is type(b) or (isinstance(a, float) and isinstance(b, float) and a == b)
    if op == "~=":
        return not _binop("==", a, b)
    if op in ("<", ">", "<=", ">="):
        if isinstance(a, str) and isinstance(b, str):
            x, y = a, b
        else:
            x, y = _num(a), _num(b)
        return {"<": x < y, ">": x > y, "<=": x <= y, ">=": x >= y}[op]
    raise LuaError("unknown operator " + op)


class _Scope:
    def __init__(self, parent, globals_):
        self.vars = {}
        self.parent = parent
        self.g = globals_

    def lookup(self, name):
        s = self
        while s is not None:
            if name in s.vars:
                return s
            s = s.parent
        return None

    def get(self, name):
        s = self.lookup(name)
        return s.vars[name] if s else self.g.get(name)

    def set(self, name, v):
        s = self.lookup(name)
        if s:
            s.vars[name] = v
        else:
            self.g.set(name, v)


class _Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k]

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def accept(self, kind, val=None):
        tok = self.peek()
        if tok[0] == kind and (val is None or tok[1] == val):
            self.i += 1
            return tok
        return None

    def expect(self, kind, val=None):
        tok = self.accept(kind, val)
        if not tok:
            raise LuaError("expected %s %s near %r" % (kind, val or "", self.peek()[1]))
        return tok

    # statements -> list of callables(scope) that may return ("return", values)
    def block(self, terminators):
        stmts = []
        while True:
            tok = self.peek()
            if tok[0] == "eof" or (tok[0] == "kw" and tok[1] in terminators):
                return stmts
            st = self.statement()
            if st is not None:
                stmts.append(st)

    def statement(self):
        tok = self.peek()
        if tok == ("op", ";"):
            self.next()
            return None
        if tok[0] == "kw":
            kw = tok[1]
            if kw == "local":
                self.next()
                if self.accept("kw", "function"):
                    name = self.expect("name")[1]
                    self._skip_function_body()
                    return lambda sc, name=name: sc.vars.__setitem__(name, LuaFunction(name))
                names = [self.expect("name")[1]]
                while self.accept("op", ","):
                    names.append(self.expect("name")[1])
                exprs = self.exprlist() if self.accept("op", "=") else []

                def run(sc, names=names, exprs=exprs):
                    vals = _eval_list(exprs, sc)
                    for k, n in enumerate(names):
                        sc.vars[n] = vals[k] if k < len(vals) else None
                return run
            if kw == "if":
                self.next()
                arms = []
                cond = self.expr()
                self.expect("kw", "then")
                body = self.block({"elseif", "else", "end"})
                arms.append((cond, body))
                els = None
                while True:
                    if self.accept("kw", "elseif"):
                        c = self.expr()
                        self.expect("kw", "then")
                        arms.append((c, self.block({"elseif", "else", "end"})))
                        continue
                    if self.accept("kw", "else"):
                        els = self.block({"end"})
                    self.expect("kw", "end")
                    break

                def run(sc, arms=arms, els=els):
                    for c, b in arms:
                        if truthy(c(sc)):
                            return _run_block(b, _Scope(sc, sc.g))
                    if els is not None:
                        return _run_block(els, _Scope(sc, sc.g))
                return run
            if kw == "function":
                self.next()
                name = self.expect("name")[1]
                while self.accept("op", ".") or self.accept("op", ":"):
                    name += "." + self.expect("name")[1]
                self._skip_function_body()
                return lambda sc, name=name: sc.set(name, LuaFunction(name)) if "." not in name else None
            if kw == "do":
                self.next()
                body = self.block({"end"})
                self.expect("kw", "end")
                return lambda sc, body=body: _run_block(body, _Scope(sc, sc.g))
            if kw == "return":
                self.next()
                exprs = [] if self.peek()[0] in ("eof",) or self.peek() == ("kw", "end") else self.exprlist()
                return lambda sc, exprs=exprs: ("return", _eval_list(exprs, sc))
            raise LuaError("unsupported statement '%s'" % kw)
        # assignment or call
        target = self.suffixed()
        if self.peek() in (("op", "="), ("op", ",")):
            targets = [target]
            while self.accept("op", ","):
                targets.append(self.suffixed())
            self.expect("op", "=")
            exprs = self.exprlist()

            def run(sc, targets=targets, exprs=exprs):
                vals = _eval_list(exprs, sc)
                for k, t in enumerate(targets):
                    t.assign(sc, vals[k] if k < len(vals) else None)
            return run
        if not target.is_call:
            raise LuaError("syntax error near %r" % (self.peek()[1],))
        return lambda sc, target=target: (target.eval_multi(sc), None)[1]

    def _skip_function_body(self):
        self.expect("op", "(")
        depth = 1
        while depth:
            tok = self.next()
            if tok == ("op", "("):
                depth += 1
            elif tok == ("op", ")"):
                depth -= 1
            elif tok[0] == "eof":
                raise LuaError("unfinished function")
        self._skip_until_end()

    def _skip_until_end(self):
        depth = 1
        while depth:
            tok = self.next()
            if tok[0] == "eof":
                raise LuaError("unfinished block")
            if tok[0] == "kw":
                if tok[1] in ("if", "do", "function", "repeat"):
                    depth += 1
                    if tok[1] == "function":
                        # its parameter list, then its body up to its own end
                        pass
                elif tok[1] in ("for", "while"):
                    # `for ... do` / `while ... do`: the `do` opens the block
                    pass
                elif tok[1] in ("end", "until"):
                    depth -= 1

    def exprlist(self):
        out = [self.expr()]
        while self.accept("op", ","):
            out.append(self.expr())
        return out

    def expr(self, limit=0):
        tok = self.peek()
        if tok in (("kw", "not"), ("op", "-"), ("op", "#")):
            self.next()
            operand = self.expr(_UNARY_PRI)
            op = tok[1]
            if op == "not":
                left = _Expr(lambda sc, e=operand: not truthy(e(sc)))
            elif op == "-":
                left = _Expr(lambda sc, e=operand: -_num(e(sc)))
            else:
                def length(sc, e=operand):
                    v = e(sc)
                    if isinstance(v, LuaTable):
                        return float(v.length())
                    if isinstance(v, str):
                        return float(len(v))
                    raise LuaError("attempt to get length of a %s value" % type(v).__name__)
                left = _Expr(length)
        else:
            left = self.simple()
        while True:
            tok = self.peek()
            op = tok[1] if tok[0] in ("op", "kw") else None
            if op not in _BINPRI or _BINPRI[op][0] <= limit:
                return left
            self.next()
            right = self.expr(_BINPRI[op][1])
            if op == "and":
                left = _Expr(lambda sc, a=left, b=right: (lambda x: b(sc) if truthy(x) else x)(a(sc)))
            elif op == "or":
                left = _Expr(lambda sc, a=left, b=right: (lambda x: x if truthy(x) else b(sc))(a(sc)))
            else:
                left = _Expr(lambda sc, a=left, b=right, op=op: _binop(op, a(sc), b(sc)))

    def simple(self):
        tok = self.peek()
        if tok[0] == "num":
            self.next()
            return _Expr(lambda sc, v=tok[1]: v)
        if tok[0] == "str":
            self.next()
            return _Expr(lambda sc, v=tok[1]: v)
        if tok[0] == "kw" and tok[1] in ("nil", "true", "false"):
            self.next()
            v = {"nil": None, "true": True, "false": False}[tok[1]]
            return _Expr(lambda sc, v=v: v)
        if tok == ("op", "{"):
            return self.table()
        if tok == ("kw", "function"):
            self.next()
            self._skip_function_body()
            return _Expr(lambda sc: LuaFunction("anonymous"))
        return self.suffixed()

    def table(self):
        self.expect("op", "{")
        items = []  # (key expr or None, value expr)
        while not self.accept("op", "}"):
            if self.peek() == ("op", "["):
                self.next()
                k = self.expr()
                self.expect("op", "]")
                self.expect("op", "=")
                items.append((k, self.expr()))
            elif self.peek()[0] == "name" and self.peek(1) == ("op", "="):
                name = self.next()[1]
                self.next()
                items.append((_Expr(lambda sc, n=name: n), self.expr()))
            else:
                items.append((None, self.expr()))
            if not (self.accept("op", ",") or self.accept("op", ";")):
                self.expect("op", "}")
                break

        def build(sc, items=items):
            t = LuaTable()
            pos = 1
            for k, v in items:
                if k is None:
                    t.set(pos, v(sc))
                    pos += 1
                else:
                    t.set(k(sc), v(sc))
            return t
        return _Expr(build)

    def suffixed(self):
        tok = self.peek()
        if tok[0] == "name":
            self.next()
            node = _Ref(name=tok[1])
        elif tok == ("op", "("):
            self.next()
            e = self.expr()
            self.expect("op", ")")
            node = _Ref(value=e)
        else:
            raise LuaError("unexpected %r" % (tok[1],))
        while True:
            tok = self.peek()
            if tok == ("op", "."):
                self.next()
                key = self.expect("name")[1]
                node = _Ref(obj=node, key=_Expr(lambda sc, k=key: k))
            elif tok == ("op", "["):
                self.next()
                k = self.expr()
                self.expect("op", "]")
                node = _Ref(obj=node, key=k)
            elif tok == ("op", ":"):
                self.next()
                meth = self.expect("name")[1]
                args = self.call_args()
                node = _Ref(call=(node, args, meth))
            elif tok == ("op", "(") or tok == ("op", "{") or tok[0] == "str":
                args = self.call_args()
                node = _Ref(call=(node, args, None))
            else:
                return node

    def call_args(self):
        tok = self.peek()
        if tok[0] == "str":
            self.next()
            return [_Expr(lambda sc, v=tok[1]: v)]
        if tok == ("op", "{"):
            return [self.table()]
        self.expect("op", "(")
        if self.accept("op", ")"):
            return []
        args = self.exprlist()
        self.expect("op", ")")
        return args


class _Expr:
    def __init__(self, fn):
        self.fn = fn

    def __call__(self, sc):
        return self.fn(sc)


class _Ref(_Expr):
    """A name, field access or call: evaluates, assigns, calls."""

    def __init__(self, name=None, obj=None, key=None, value=None, call=None):
        self.name, self.obj, self.key, self.value, self.callspec = name, obj, key, value, call
        self.is_call = call is not None
        super().__init__(self._eval)

    def _eval(self, sc):
        r = self.eval_multi(sc)
        return r[0] if isinstance(r, list) else r

    def eval_multi(self, sc):
        if self.callspec is not None:
            fnode, args, meth = self.callspec
            f = fnode(sc)
            argv = _eval_list(args, sc)
            if meth is not None:
                if not isinstance(f, LuaTable):
                    raise LuaError("attempt to call method on a non-table")
                argv = [f] + argv
                f = f.get(meth)
            if not isinstance(f, LuaFunction):
                raise LuaError("attempt to call a %s value" % ("nil" if f is None else type(f).__name__))
            r = f(argv)
            return r if isinstance(r, list) else [r]
        if self.name is not None:
            return sc.get(self.name)
        if self.value is not None:
            return self.value(sc)
        t = self.obj(sc)
        if not isinstance(t, LuaTable):
            raise LuaError("attempt to index a %s value" % ("nil" if t is None else type(t).__name__))
        return t.get(self.key(sc))

    def assign(self, sc, v):
        if self.name is not None:
            sc.set(self.name, v)
            return
        if self.obj is None:
            raise LuaError("cannot assign to this expression")
        t = self.obj(sc)
        if not isinstance(t, LuaTable):
            raise LuaError("attempt to index a %s value" % ("nil" if t is None else type(t).__name__))
        t.set(self.key(sc), v)


def _eval_list(exprs, sc):
    out = []
    for k, e in enumerate(exprs):
        if k == len(exprs) - 1 and isinstance(e, _Ref) and e.is_call:
            out.extend(e.eval_multi(sc))
        else:
            out.append(e(sc))
    return out


def _run_block(stmts, sc):
    for st in stmts:
        r = st(sc)
        if isinstance(r, tuple) and r and r[0] == "return":
            return r
    return None


# ---------------------------------------------------------------- environment
def _math_table():
    def f1(fn):
        return LuaFunction("math", lambda a: [float(fn(_num(a[0])))])
    t = LuaTable(hsh={
        "pi": math.pi, "huge": math.inf,
        "sqrt": f1(math.sqrt), "exp": f1(math.exp), "log": LuaFunction("math.log", lambda a: [
            math.log(_num(a[0])) if len(a) < 2 or a[1] is None else math.log(_num(a[0]), _num(a[1]))]),
        "sin": f1(math.sin), "cos": f1(math.cos), "tan": f1(math.tan), "abs": f1(abs),
        "floor": f1(math.floor), "ceil": f1(math.ceil), "rad": f1(math.radians), "deg": f1(math.degrees),
        "atan": f1(math.atan), "asin": f1(math.asin), "acos": f1(math.acos),
        "max": LuaFunction("math.max", lambda a: [max(_num(x) for x in a)]),
        "min": LuaFunction("math.min", lambda a: [min(_num(x) for x in a)]),
    })
    return t


def evaluate(text, initconds=None):
    """Run config text; returns the global table (LuaTable).  `initconds`
    maps initial-condition names to callables setup(globals, args) that
    may modify the globals (initconds.setup_in_lua)."""
    g = LuaTable()
    g.set("math", _math_table())
    g.set("package", LuaTable(hsh={"path": ""}))
    g.set("require", LuaFunction("require", lambda a: [None]))
    g.set("print", LuaFunction("print", lambda a: [None]))
    g.set("tostring", LuaFunction("tostring", lambda a: [_tostr(a[0])]))
    g.set("tonumber", LuaFunction("tonumber", lambda a: [_num(a[0])]))
    chosen = []
    ic = LuaTable()
    if initconds:
        for name, fn in initconds.items():
            def setup(args, name=name, fn=fn):
                chosen.append(name)
                fn(g, args[0] if args else None)
                return [None]
            ic.set(name, LuaTable(hsh={"name": name, "setup": LuaFunction("setup", setup)}))
    g.set("initConds", ic)
    stmts = _Parser(tokenize(text)).block(set())
    _run_block(stmts, _Scope(None, g))
    g.set("__initCondNames", LuaTable(arr=list(chosen)))
    return g


def _get(g, key, default=None):
    v = g.get(key)
    return default if v is None else v


def load(path_or_text, precision="fp64", flags=_config.FLAG_REFERENCE, t_end=0.0):
    """Read a config.lua-style file (or its text).  Returns (desc, info)
    where desc is the hydro_sim_desc the reference would build
    (src/HydroGPUApp.cpp:164-230; the limiter key is `slopeLimiter`,
    src/Solver/Solver.cpp:285-287, default Superbee) and info holds
    initCondName, the initial-state function (initconds) and the keys this
    build does not act on (maxFrames, ...); useGravity and
    gaussSeidelMaxIter go into the descriptor."""
    from . import initconds

    text = path_or_text
    base_dir = os.getcwd()
    if "\n" not in path_or_text and path_or_text.endswith(".lua"):
        with open(path_or_text) as f:
            text = f.read()
        base_dir = os.path.dirname(os.path.abspath(path_or_text))
    g = evaluate(text, initconds.lua_setups())
    size = to_py(g.get("size"))
    if not size:
        raise LuaError("couldn't find any size information")
    size = [int(v) for v in (size if isinstance(size, list) else [size[k] for k in sorted(size)])]
    # store dimension as last non-1 size (src/HydroGPUApp.cpp:208-215)
    while len(size) > 1 and size[-1] == 1:
        size.pop()
    xmin = [float(v) for v in to_py(_get(g, "xmin", LuaTable(arr=[-.5, -.5, -.5])))]
    xmax = [float(v) for v in to_py(_get(g, "xmax", LuaTable(arr=[.5, .5, .5])))]
    bm = to_py(g.get("boundaryMethods")) or []
    bms = []
    for a in range(3):
        e = bm[a] if isinstance(bm, list) and a < len(bm) else None
        bms.append({"min": e.get("min", "FREEFLOW"), "max": e.get("max", "FREEFLOW")} if isinstance(e, dict)
                   else {"min": "FREEFLOW", "max": "FREEFLOW"})
    defs = to_py(g.get("defs")) or {}
    defs = {k: float(v) for k, v in defs.items() if isinstance(v, (int, float)) and not isinstance(v, bool)}
    solver = _get(g, "solverName", "EulerRoe")
    integ = _get(g, "integratorName", "ForwardEuler")
    lim = _get(g, "slopeLimiter", "Superbee")
    cfl = g.get("cfl")
    if cfl is None:
        cfl = 0.5  # HydroGPUApp's own default (src/HydroGPUApp.cpp:74)
    desc = _config.make_desc(size, xmin=xmin, xmax=xmax, solverName=solver, integratorName=integ,
                             slopeLimiter=lim, boundaryMethods=bms, cfl=cfl,
                             useFixedDT=truthy(g.get("useFixedDT")), fixedDT=float(_get(g, "fixedDT", 0.001)),
                             defs=defs, precision=precision, flags=flags, t_end=t_end,
                             useGravity=truthy(g.get("useGravity")),
                             gaussSeidelMaxIter=int(_get(g, "gaussSeidelMaxIter", 20)))
    names = to_py(g.get("__initCondNames")) or []
    name = names[-1] if names else None
    info = {
        "initCondName": name,
        "initState": initconds.state_function(name, defs.get("idealGas_heatCapacityRatio")) if name else None,
        "initChannels": initconds.channels_function(name, defs.get("idealGas_heatCapacityRatio")) if name else None,
        "solidFilename": g.get("solidFilename"),
        "slopeLimiterName": g.get("slopeLimiterName"),
        "maxFrames": g.get("maxFrames"),
        "useGravity": truthy(g.get("useGravity")),
        "custom_initState": isinstance(g.get("initState"), LuaFunction) and name is None,
        "calcSolid": isinstance(g.get("calcSolid"), LuaFunction),
        "globals": g,
    }
    if info["calcSolid"]:
        raise LuaError("calcSolid (a Lua function of the cell position) is not evaluated here; pass the solid "
                       "cells as an array (Sim.set_solid / hydro_sim_set_solid)")
    sf = info["solidFilename"]
    if isinstance(sf, str) and info["initChannels"] is not None:
        # SelfGravitation::resetState replaces the initial condition's solid
        # channel by the image (src/Solver/SelfGravitation.cpp:56-74)
        from . import io as _io

        path = sf if os.path.isabs(sf) else os.path.join(base_dir, sf)
        channels = info["initChannels"]

        def with_image(d, channels=channels, path=path):
            q, pot, _ = channels(d)
            return q, pot, _io.solid_from_image(d, _io.read_png(path))
        info["initChannels"] = with_image
    return desc, info
