"""Reader for the subset of the TDB (Thermo-Calc database) format the
N-component engine needs, in place of pycalphad's ``Database`` +
``Model.redlich_kister_sum`` + ``ideal_mixing_energy`` (``simulation.py:196-214``,
``ppf_utils.py:132-160`` of the reference).

Supported: ``ELEMENT``, ``FUNCTION``, ``PHASE`` with one sublattice,
``CONSTITUENT``, ``PARAMETER G(phase, A; 0)`` and binary
``PARAMETER L(phase, A, B; k)``; piecewise temperature ranges (``T0 expr; T1 Y
expr; ... Tn N``); expressions built from numbers, ``T``, ``LN(T)``, integer
powers ``T**n``, ``+ - * /``, parentheses and references to FUNCTIONs (with or
without the trailing ``#``).  Other commands (``TYPE_DEFINITION``,
``LIST_OF_REFERENCES``, ``DATABASE_INFO``, ...) are skipped.  Comment lines start
with ``$``; commands end with ``!``.

Every temperature function becomes a :class:`Piecewise` of
:class:`Poly` terms c * T**p * (ln T)**q.  Outside a function's outermost
ranges the nearest range is used (pycalphad returns 0 there, which no
simulation can use).
"""

import math
import re

import numpy as np

# device basis of ppf_tpoly, in order: (power of T, power of ln T)
BASIS = [(0, 0), (1, 0), (1, 1), (2, 0), (3, 0), (-1, 0), (7, 0), (-9, 0)]
BASIS_INDEX = {k: i for i, k in enumerate(BASIS)}


class TDBError(ValueError):
    pass


class Poly:
    """sum of c * T**p * ln(T)**q, stored as {(p, q): c}"""

    def __init__(self, terms=None):
        self.terms = {k: v for k, v in (terms or {}).items() if v != 0.0}

    @staticmethod
    def const(c):
        return Poly({(0, 0): float(c)})

    def is_const(self):
        return all(k == (0, 0) for k in self.terms)

    def value_const(self):
        return self.terms.get((0, 0), 0.0)

    def __add__(self, o):
        t = dict(self.terms)
        for k, v in o.terms.items():
            t[k] = t.get(k, 0.0) + v
        return Poly(t)

    def __neg__(self):
        return Poly({k: -v for k, v in self.terms.items()})

    def __mul__(self, o):
        t = {}
        for (p1, q1), v1 in self.terms.items():
            for (p2, q2), v2 in o.terms.items():
                k = (p1 + p2, q1 + q2)
                t[k] = t.get(k, 0.0) + v1 * v2
        return Poly(t)

    def power(self, n):
        if self.is_const():
            return Poly.const(self.value_const() ** n)
        if len(self.terms) == 1:
            (p, q), v = next(iter(self.terms.items()))
            if float(n).is_integer():
                n = int(n)
                return Poly({(p * n, q * n): v ** n})
        raise TDBError("unsupported power of a temperature expression")

    def evaluate(self, T):
        T = np.asarray(T, dtype=np.float64)
        out = np.zeros_like(T)
        lnT = np.log(T)
        for (p, q), v in self.terms.items():
            out = out + v * T ** p * lnT ** q
        return out


class Piecewise:
    """Ranges [(T_hi, Poly)] in increasing T_hi; range r covers
    (T_hi[r-1], T_hi[r]] (the first from T_lo).  Constants use one range."""

    def __init__(self, T_lo, ranges):
        self.T_lo = float(T_lo)
        self.ranges = [(float(h), p) for h, p in ranges]

    @staticmethod
    def const(c):
        return Piecewise(-math.inf, [(math.inf, Poly.const(c))])

    @property
    def breakpoints(self):
        return [h for h, _ in self.ranges]

    def poly_at(self, T):
        for h, p in self.ranges:
            if T <= h:
                return p
        return self.ranges[-1][1]

    def combine(self, other, op):
        """pointwise op over the union of both breakpoint sets"""
        bps = sorted(set(self.breakpoints) | set(other.breakpoints))
        lo = max(self.T_lo, other.T_lo) if math.isfinite(self.T_lo) and math.isfinite(other.T_lo) else \
            (self.T_lo if math.isfinite(self.T_lo) else other.T_lo)
        out = []
        prev = lo
        for h in bps:
            mid = _midpoint(prev, h)
            out.append((h, op(self.poly_at(mid), other.poly_at(mid))))
            prev = h
        return Piecewise(lo, _merge_equal(out))

    def __add__(self, o):
        return self.combine(o, lambda a, b: a + b)

    def __sub__(self, o):
        return self.combine(o, lambda a, b: a + (-b))

    def __mul__(self, o):
        return self.combine(o, lambda a, b: a * b)

    def __neg__(self):
        return Piecewise(self.T_lo, [(h, -p) for h, p in self.ranges])

    def is_const(self):
        return all(p.is_const() for _, p in self.ranges)

    def value_const(self):
        vals = {p.value_const() for _, p in self.ranges}
        if len(vals) != 1:
            raise TDBError("not a constant")
        return vals.pop()

    def evaluate(self, T):
        """FP64 value at T (array); the nearest range outside [T_lo, T_hi]."""
        T = np.asarray(T, dtype=np.float64)
        out = np.zeros_like(T)
        done = np.zeros(T.shape, dtype=bool)
        for h, p in self.ranges:
            sel = (T <= h) & ~done
            if sel.any():
                out = np.where(sel, p.evaluate(np.where(sel, T, 1.0)), out)
            done |= sel
        if (~done).any():
            p = self.ranges[-1][1]
            out = np.where(done, out, p.evaluate(np.where(done, 1.0, T)))
        return out

    def device_table(self, max_ranges):
        """(T_hi[max_ranges], coeff[max_ranges][len(BASIS)], nranges) for ppf_tpoly."""
        if len(self.ranges) > max_ranges:
            raise TDBError("a temperature function has %d ranges; at most %d are supported"
                           % (len(self.ranges), max_ranges))
        T_hi = np.full(max_ranges, np.inf)
        c = np.zeros((max_ranges, len(BASIS)))
        for r, (h, p) in enumerate(self.ranges):
            T_hi[r] = h
            for k, v in p.terms.items():
                if k not in BASIS_INDEX:
                    raise TDBError("term T**%d * ln(T)**%d is outside the supported basis "
                                   "(1, T, T ln T, T^2, T^3, 1/T, T^7, T^-9)" % k)
                c[r, BASIS_INDEX[k]] = v
        return T_hi, c, len(self.ranges)


def _midpoint(a, b):
    if not math.isfinite(a) and not math.isfinite(b):
        return 1000.0
    if not math.isfinite(a):
        return b - 1.0
    if not math.isfinite(b):
        return a + 1.0
    return 0.5 * (a + b)


def _merge_equal(ranges):
    out = []
    for h, p in ranges:
        if out and out[-1][1].terms == p.terms:
            out[-1] = (h, p)
        else:
            out.append((h, p))
    return out


# ------------------------------------------------------------ expressions --
_TOKEN = re.compile(r"\s*(?:(?P<num>(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)|(?P<name>[A-Za-z_][A-Za-z0-9_]*#?)"
                    r"|(?P<op>\*\*|[-+*/(),]))")


def _tokens(text):
    pos = 0
    out = []
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise TDBError("cannot parse expression near %r" % text[pos:pos + 20])
        if m.group("num") is not None:
            out.append(("num", float(m.group("num"))))
        elif m.group("name") is not None:
            out.append(("name", m.group("name").rstrip("#").upper()))
        else:
            out.append(("op", m.group("op")))
        pos = m.end()
        while pos < len(text) and text[pos].isspace():
            pos += 1
    return out


class _Parser:
    def __init__(self, tokens, resolve):
        self.t = tokens
        self.i = 0
        self.resolve = resolve

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, kind=None, val=None):
        tok = self.peek()
        if tok[0] is None or (kind and tok[0] != kind) or (val and tok[1] != val):
            raise TDBError("unexpected token %r" % (tok,))
        self.i += 1
        return tok

    def expr(self):
        v = self.term()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            r = self.term()
            v = v + r if op == "+" else v - r
        return v

    def term(self):
        v = self.factor()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            r = self.factor()
            if op == "*":
                v = v * r
            else:
                if not r.is_const():
                    raise TDBError("division by a temperature expression")
                v = v * Piecewise.const(1.0 / r.value_const())
        return v

    def factor(self):
        base = self.unary()
        if self.peek() == ("op", "**"):
            self.take()
            ex = self.unary()
            if not ex.is_const():
                raise TDBError("non-constant exponent")
            n = ex.value_const()
            return Piecewise(base.T_lo, [(h, p.power(n)) for h, p in base.ranges])
        return base

    def unary(self):
        if self.peek() == ("op", "-"):
            self.take()
            return -self.unary()
        if self.peek() == ("op", "+"):
            self.take()
            return self.unary()
        return self.primary()

    def primary(self):
        kind, val = self.peek()
        if kind == "num":
            self.take()
            return Piecewise.const(val)
        if kind == "op" and val == "(":
            self.take()
            v = self.expr()
            self.take("op", ")")
            return v
        if kind == "name":
            self.take()
            if val == "T":
                return Piecewise(-math.inf, [(math.inf, Poly({(1, 0): 1.0}))])
            if val in ("LN", "LOG"):
                self.take("op", "(")
                arg = self.expr()
                self.take("op", ")")
                if not (len(arg.ranges) == 1 and arg.ranges[0][1].terms == {(1, 0): 1.0}):
                    raise TDBError("LN() is supported of T only")
                return Piecewise(-math.inf, [(math.inf, Poly({(0, 1): 1.0}))])
            return self.resolve(val)
        raise TDBError("unexpected token %r" % ((kind, val),))


def parse_piecewise(text, resolve):
    """'T0 expr; T1 Y expr; ... Tn N [ref]' -> Piecewise"""
    parts = [p.strip() for p in text.split(";")]
    head = parts[0].split(None, 1)
    if len(head) < 2:
        raise TDBError("temperature function without an expression: %r" % text)
    T_lo = float(head[0])
    exprs = [head[1]]
    highs = []
    for part in parts[1:]:
        if not part:
            continue
        fields = part.split()
        highs.append(float(fields[0]))
        flag = fields[1].upper() if len(fields) > 1 else "N"
        if flag == "Y":
            rest = part.split(None, 2)
            exprs.append(rest[2] if len(rest) > 2 else "")
        else:
            break
    if len(highs) != len(exprs):
        raise TDBError("unbalanced temperature ranges in %r" % text)
    ranges = []
    for h, ex in zip(highs, exprs):
        pw = _Parser(_tokens(ex), resolve).expr()
        if not pw.ranges:
            raise TDBError("empty expression")
        # an expression referencing other functions keeps their breakpoints,
        # clipped to this range
        lo = ranges[-1][0] if ranges else T_lo
        for bh, p in pw.ranges:
            if bh <= lo:
                continue
            ranges.append((min(bh, h), p))
            if bh >= h:
                break
        if not ranges or ranges[-1][0] < h:
            ranges.append((h, pw.poly_at(h)))
    return Piecewise(T_lo, _merge_equal(ranges))


# --------------------------------------------------------------- database --
class Database:
    """elements, functions {name: Piecewise}, phases {name: constituents},
    parameters {(kind 'G'|'L', phase, constituents tuple, order): Piecewise}.

    Read leniently, as a database of many elements is used for a few: what
    is outside the supported subset is recorded, not refused, and refused
    only when a model for the components that touch it is built
    (``thermo.load_tdb_model``):
      * phases of several sublattices (``multi_sublattice``, name -> count)
        and their parameters (``unsupported``);
      * functions and parameters whose expressions fall outside the
        temperature basis (``unsupported``, with the reason);
      * parameter types other than G and L (TC, BMAGN, ...) are skipped
        (``ignored``): the reference's phase model is
        ``Model.redlich_kister_sum`` over G and L parameters plus the ideal
        mixing energy (simulation.py:206-214), which leaves out the magnetic
        and other contributions these describe.
    A G parameter of several interacting species is an interaction term like
    L with the same constituents (redlich_kister_sum takes both types
    alike); it is stored under kind 'L'."""

    def __init__(self, text):
        self.elements = []
        self.functions = {}
        self._function_text = {}
        self._function_error = {}
        self.phases = {}
        self.multi_sublattice = {}
        self.parameters = {}
        self.unsupported = {}
        self.ignored = {}
        commands = _commands(text)
        # functions may reference functions defined later: resolve lazily
        for cmd in commands:
            word = cmd.split(None, 1)[0].upper()
            if word.startswith("FUNCT"):
                name, rest = cmd.split(None, 2)[1:]
                self._function_text[name.upper().rstrip("#")] = rest
        for name in list(self._function_text):
            try:
                self._function(name)
            except TDBError as e:
                self._function_error[name] = e
                self.functions.pop(name, None)
        for cmd in commands:
            word = cmd.split(None, 1)[0].upper()
            if word.startswith("ELEM"):
                el = cmd.split()[1].upper()
                if el not in ("VA", "/-"):
                    self.elements.append(el)
            elif word.startswith("PHASE"):
                f = cmd.split()
                name = f[1].upper().split(":")[0]
                nsub = int(f[3])
                if nsub != 1:
                    self.multi_sublattice[name] = nsub
                    continue
                self.phases.setdefault(name, [])
            elif word.startswith("CONST"):
                m = re.match(r"\S+\s+([^\s:]+)\s*:(.*):", cmd, re.S)
                if not m:
                    raise TDBError("cannot read %r" % cmd)
                name = m.group(1).upper().split(":")[0]
                subl = [s for s in m.group(2).split(":") if s.strip()]
                if name in self.multi_sublattice or len(subl) != 1:
                    self.multi_sublattice.setdefault(name, len(subl))
                    self.phases.pop(name, None)
                    continue
                self.phases[name] = [c.strip().upper().rstrip("%") for c in subl[0].split(",") if c.strip()]
            elif word.startswith("PARA"):
                self._parameter(cmd)

    def _function(self, name):
        name = name.upper().rstrip("#")
        if name in self._function_error:
            raise self._function_error[name]
        if name in self.functions:
            return self.functions[name]
        if name not in self._function_text:
            raise TDBError("undefined function %s" % name)
        self.functions[name] = None  # recursion guard
        self.functions[name] = parse_piecewise(self._function_text[name], self._resolve)
        return self.functions[name]

    def _resolve(self, name):
        if name in self.functions and self.functions[name] is None:
            raise TDBError("recursive function %s" % name)
        return self._function(name)

    def _parameter(self, cmd):
        m = re.match(r"\S+\s+(\w+)\s*\(\s*([^,]+?)\s*,\s*([^;]+?)\s*;\s*(\d+)\s*\)\s*(.*)$", cmd, re.S | re.I)
        if not m:
            raise TDBError("cannot read parameter %r" % cmd[:60])
        kind = m.group(1).upper()
        phase = m.group(2).upper()
        cons_text = m.group(3).upper()
        order = int(m.group(4))
        if kind not in ("G", "L"):
            self.ignored[kind] = self.ignored.get(kind, 0) + 1
            return
        if ":" in cons_text:
            self.unsupported[(phase, cons_text, order)] = "parameter of several sublattices"
            return
        cons = tuple(c.strip().rstrip("%") for c in cons_text.split(","))
        key = ("G" if len(cons) == 1 else "L", phase, cons, order)
        try:
            f = parse_piecewise(m.group(5), self._resolve)
        except TDBError as e:
            self.unsupported[(phase, ",".join(cons), order)] = str(e)
            return
        if key in self.parameters:
            raise TDBError("parameter %s(%s,%s;%d) given twice" % (key[0], phase, ",".join(cons), order))
        self.parameters[key] = f

    def constant(self, name):
        """value of a constant FUNCTION (the engine parameters, NCGPU_new.py:358-383)"""
        f = self._function(name)
        if not f.is_const():
            raise TDBError("function %s is not a constant" % name)
        return f.value_const()


def _commands(text):
    lines = [ln for ln in text.splitlines() if not ln.lstrip().startswith("$")]
    body = "\n".join(lines)
    out = []
    for cmd in body.split("!"):
        cmd = " ".join(cmd.split())
        if cmd:
            out.append(cmd)
    return out


def read(path):
    with open(path) as f:
        return Database(f.read())
