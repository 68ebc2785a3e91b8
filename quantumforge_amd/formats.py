"""Circuit formats: OpenQASM 2.0 import/export and the JSON SerializedCircuit
form, with the crate's behaviour (src/qasm.rs, src/qasm_export.rs,
src/serialize.rs).

QASM import follows the crate's statement-by-statement reader: `OPENQASM`
header required, `include` skipped, several `;`-separated statements per line,
`qreg`/`creg` registers laid out in declaration order, `gate` definitions
expanded inline, `barrier` ignored, `reset` rejected, `measure q -> c`.  Gate
lowerings are the crate's: u1(l) -> Rz(l); u2(p, l) -> Rz(p) Ry(pi/2) Rz(l);
u3(t, p, l) -> Rz(p) Ry(t) Rz(l); ccx -> the 15-gate Clifford+T circuit.  A
whole-register operand names the register's first qubit.  Parameter
expressions take numbers, `pi`, `e`, + - * / and parentheses, evaluated after
the crate's textual substitution of `pi` and then of every `e`; parameters
at user-gate call sites are plain numbers.
"""
from __future__ import annotations

import json
import math
from typing import Dict, List, Optional, Tuple

from .core import Circuit, Gate, NoiseChannel, _GATE_NAMES

__all__ = ["QasmError", "parse_qasm", "export_qasm", "export_qasm_with_qubits", "to_json", "to_json_bytes",
           "from_json", "from_json_bytes"]


class QasmError(ValueError):
    """QasmError, src/qasm.rs:27-51: the line and a message."""

    def __init__(self, line: int, message: str) -> None:
        self.line = int(line)
        self.message = message
        super().__init__(f"QASM error at line {self.line}: {message}")


_PARAM_GATES = {"rx", "Rx", "ry", "Ry", "rz", "Rz", "u1", "U1", "u2", "U2", "u3", "U3"}


def _split_name(s: str) -> str:
    for i, ch in enumerate(s):
        if ch in "( [":
            return s[:i]
    return s


# ---------------------------------------------------------------------------
# Parameter expressions (src/qasm.rs:913-1077)
# ---------------------------------------------------------------------------
def _tokens(s: str) -> Optional[list]:
    out, i = [], 0
    while i < len(s):
        ch = s[i]
        if ch in " \t":
            i += 1
        elif ch in "+-*/()":
            out.append(ch)
            i += 1
        elif ch.isdigit() or ch == ".":
            j = i
            while j < len(s) and (s[j].isdigit() or s[j] == "."):
                j += 1
            try:
                out.append(float(s[i:j]))
            except ValueError:
                return None
            i = j
        else:
            return None
    return out


def _eval(tokens: list) -> Optional[float]:
    pos = [0]

    def factor():
        if pos[0] >= len(tokens):
            return None
        t = tokens[pos[0]]
        if isinstance(t, float):
            pos[0] += 1
            return t
        if t in "+-":
            pos[0] += 1
            v = factor()
            return None if v is None else (-v if t == "-" else v)
        if t == "(":
            pos[0] += 1
            v = expr()
            if v is None or pos[0] >= len(tokens) or tokens[pos[0]] != ")":
                return None
            pos[0] += 1
            return v
        return None

    def term():
        v = factor()
        while v is not None and pos[0] < len(tokens) and tokens[pos[0]] in ("*", "/"):
            op = tokens[pos[0]]
            pos[0] += 1
            r = factor()
            if r is None or (op == "/" and r == 0.0):
                return None
            v = v * r if op == "*" else v / r
        return v

    def expr():
        v = term()
        while v is not None and pos[0] < len(tokens) and tokens[pos[0]] in ("+", "-"):
            op = tokens[pos[0]]
            pos[0] += 1
            r = term()
            if r is None:
                return None
            v = v + r if op == "+" else v - r
        return v

    v = expr()
    return v if v is not None and pos[0] == len(tokens) else None


def _rust_float(x: float) -> str:
    # Rust's Display of f64 (shortest representation that round-trips).
    r = repr(float(x))
    if "e" in r or "E" in r:
        m, e = r.split("e")
        e = int(e)
        digits = m.replace("-", "").replace(".", "")
        neg = m.startswith("-")
        point = (m.replace("-", "").index(".") if "." in m else len(m.replace("-", ""))) + e
        if point <= 0:
            out = "0." + "0" * (-point) + digits.rstrip("0")
        elif point >= len(digits):
            out = digits + "0" * (point - len(digits))
        else:
            out = digits[:point] + "." + digits[point:]
        r = ("-" if neg else "") + out
    if r.endswith(".0"):
        r = r[:-2]
    return r


def eval_qasm_expr(s: str) -> Optional[float]:
    """eval_qasm_expr, src/qasm.rs:913-925 (substitutes `pi`, then `e`)."""
    s = s.strip()
    if not s:
        return None
    s = s.replace("pi", _rust_float(math.pi))
    s = s.replace("e", _rust_float(math.e))
    t = _tokens(s)
    return None if t is None else _eval(t)


# ---------------------------------------------------------------------------
# Import
# ---------------------------------------------------------------------------
class _Parser:
    def __init__(self, source: str) -> None:
        self.lines = source.splitlines()
        self.pending: List[str] = []
        self.pos = 0
        self.c = Circuit()
        self.qregs: Dict[str, Tuple[int, int]] = {}
        self.cregs: Dict[str, Tuple[int, int]] = {}
        self.nq = 0
        self.nc = 0
        self.defs: Dict[str, tuple] = {}

    def err(self, msg: str) -> QasmError:
        return QasmError(self.pos, msg)

    def parse(self) -> Circuit:
        header = False
        while self.pending or self.pos < len(self.lines):
            if self.pending:
                raw = self.pending.pop(0)
            else:
                raw = self.lines[self.pos].strip()
                self.pos += 1
            if not raw or raw.startswith("//"):
                continue
            if raw.startswith("/*") and raw.endswith("*/"):
                continue
            if raw.startswith("OPENQASM"):
                header = True
                continue
            if raw.startswith("include"):
                continue
            if not header:
                raise self.err("missing OPENQASM 2.0 header")
            if "{" in raw or "}" in raw:
                line = raw
            else:
                segs = [x.strip() for x in raw.split(";") if x.strip()]
                self.pending = segs[1:] + self.pending
                line = segs[0] if segs else ""
            if not line or line.startswith("//"):
                continue
            if line.startswith("qreg"):
                self.reg(line, "qreg")
            elif line.startswith("creg"):
                self.reg(line, "creg")
            elif line.startswith("gate"):
                self.gate_def(line)
            elif line.startswith("barrier"):
                continue
            elif line.startswith("reset"):
                raise self.err("reset is not supported in pure-state simulation (requires measure+conditional-X; "
                               "use shot execution mode)")
            elif line.startswith("measure"):
                self.measure(line)
            else:
                self.application(line)
        return self.c

    def reg(self, line: str, kind: str) -> None:
        rest = line.rstrip(";").strip()[len(kind):]
        try:
            name = rest[: rest.index("[")].strip()
            size = int(rest[rest.index("[") + 1: rest.index("]")].strip())
            if size < 0:
                raise ValueError
        except ValueError:
            raise self.err(f"invalid {kind} declaration") from None
        if kind == "qreg":
            self.qregs[name] = (self.nq, size)
            self.nq += size
        else:
            self.cregs[name] = (self.nc, size)
            self.nc += size

    def gate_def(self, first: str) -> None:
        header = first.strip()
        body: List[str] = []
        if "{" in header:
            k = header.index("{")
            header, rest = header[:k], header[k + 1:]
            if rest.strip():
                body.append(rest)
        else:
            while self.pos < len(self.lines):
                nxt = self.lines[self.pos].strip()
                self.pos += 1
                if not nxt or nxt.startswith("//"):
                    continue
                if "{" in nxt:
                    k = nxt.index("{")
                    header = f"{header} {nxt[:k]}"
                    if nxt[k + 1:].strip():
                        body.append(nxt[k + 1:])
                    break
                header = f"{header} {nxt}"
        depth = 0
        while self.pos < len(self.lines):
            t = self.lines[self.pos].strip()
            self.pos += 1
            if not t or t.startswith("//"):
                continue
            for ch in t:
                if ch == "{":
                    depth += 1
                elif ch == "}":
                    if depth == 0:
                        before = t[: t.index("}")]
                        if before.strip():
                            body.append(before)
                        self.define(header, body)
                        return
                    depth -= 1
            body.append(t)
        raise self.err("unclosed gate definition (missing '}')")

    def define(self, header: str, body_lines: List[str]) -> None:
        h = header.strip()
        if h.startswith("gate "):
            h = h[5:].strip()
        name_end = h.find("(") if "(" in h else (h.find(" ") if " " in h else len(h))
        name = h[:name_end].strip()
        params: List[str] = []
        if "(" in h:
            if ")" not in h:
                raise self.err("unclosed parenthesis in gate definition")
            params = [p.strip() for p in h[h.index("(") + 1: h.index(")")].split(",") if p.strip()]
            qsrc = h[h.index(")") + 1:]
        else:
            qsrc = h[name_end:]
        qargs = qsrc.replace(",", " ").split()
        body = []
        for stmt in " ".join(body_lines).split(";"):
            stmt = stmt.strip()
            if not stmt:
                continue
            gname = _split_name(stmt)
            bparams: List[str] = []
            if "(" in stmt:
                if ")" not in stmt:
                    raise self.err("unclosed parenthesis in gate body")
                bparams = [p.strip() for p in stmt[stmt.index("(") + 1: stmt.index(")")].split(",") if p.strip()]
                qs = stmt[stmt.index(")") + 1:]
            else:
                qs = stmt[len(gname):]
            body.append((gname, bparams, [x.strip() for x in qs.split(",") if x.strip()]))
        self.defs[name] = (params, qargs, body)

    def measure(self, line: str) -> None:
        rest = line.rstrip(";").strip()[len("measure"):]
        parts = rest.split("->")
        if len(parts) != 2:
            raise self.err("measure requires 'q[i] -> c[j]'")
        self.c.measure(self.qubit(parts[0].strip()), self.cbit(parts[1].strip()))

    def call(self, line: str) -> Tuple[str, List[str]]:
        line = line.strip()
        name = _split_name(line)
        if "(" in line:
            if ")" not in line:
                raise self.err("unclosed parenthesis in gate call")
            qs = line[line.index(")") + 1:]
        else:
            qs = line[len(name):]
        return name, [x.strip().rstrip(";").strip() for x in qs.split(",") if x.strip().rstrip(";").strip()]

    def application(self, line: str) -> None:
        line = line.rstrip(";").strip()
        name = _split_name(line)
        if name in self.defs:
            self.user_gate(line, name)
        elif name in _PARAM_GATES:
            params: List[float] = []
            if "(" in line:
                if ")" not in line:
                    raise self.err("unclosed parenthesis in gate call")
                for p in line[line.index("(") + 1: line.index(")")].split(","):
                    p = p.strip()
                    if p:
                        v = eval_qasm_expr(p)
                        if v is None:
                            raise self.err(f"invalid parameter '{p}'")
                        params.append(v)
            qs = line[line.index(")") + 1:] if ")" in line else line[len(name):]
            idx = [self.qubit(x.strip().rstrip(";").strip()) for x in qs.split(",")
                   if x.strip().rstrip(";").strip()]
            self.param_gate(name, params, idx)
        else:
            n, qargs = self.call(line)
            self.builtin(n, [self.qubit(q) for q in qargs])

    def user_gate(self, line: str, name: str) -> None:
        params, formal, body = self.defs[name]
        _, actual = self.call(line)
        if len(actual) != len(formal):
            raise self.err(f"gate '{name}' expects {len(formal)} qargs, got {len(actual)}")
        amap = {f: self.qubit(a) for f, a in zip(formal, actual)}
        values: List[float] = []
        if "(" in line and ")" in line:
            for p in line[line.index("(") + 1: line.index(")")].split(","):
                p = p.strip()
                try:
                    values.append(float(p))
                except ValueError:
                    pass
        pmap = {pn: values[i] for i, pn in enumerate(params) if i < len(values)}
        for gname, bparams, bq in body:
            idx = [amap.get(x, 0) for x in bq]
            vals = []
            for bp in bparams:
                if bp in pmap:
                    vals.append(pmap[bp])
                else:
                    try:
                        vals.append(float(bp))
                    except ValueError:
                        vals.append(0.0)
            if gname in _PARAM_GATES:
                self.param_gate(gname, vals, idx)
            else:
                self.builtin(gname, idx)

    def param_gate(self, name: str, p: List[float], idx: List[int]) -> None:
        c, q = self.c, idx[0]
        if name in ("rx", "Rx"):
            c.push(Gate.Rx(p[0]), q)
        elif name in ("ry", "Ry"):
            c.push(Gate.Ry(p[0]), q)
        elif name in ("rz", "Rz", "u1", "U1"):
            c.push(Gate.Rz(p[0]), q)
        elif name in ("u2", "U2"):
            c.push(Gate.Rz(p[0]), q)
            c.push(Gate.Ry(math.pi / 2), q)
            c.push(Gate.Rz(p[1]), q)
        else:  # u3
            c.push(Gate.Rz(p[1]), q)
            c.push(Gate.Ry(p[0]), q)
            c.push(Gate.Rz(p[2]), q)

    def builtin(self, name: str, idx: List[int]) -> None:
        c = self.c
        one = {"x": "X", "X": "X", "y": "Y", "Y": "Y", "z": "Z", "Z": "Z", "h": "H", "H": "H", "s": "S", "S": "S",
               "sdg": "Sdg", "Sdg": "Sdg", "t": "T", "T": "T", "tdg": "Tdg", "Tdg": "Tdg", "id": "I", "I": "I",
               "Identity": "I"}
        if name in one:
            c.push(getattr(Gate, one[name]), idx[0])
        elif name in ("cx", "CX", "cnot", "CNOT"):
            c.cnot(idx[0], idx[1])
        elif name in ("cz", "CZ"):
            c.cz(idx[0], idx[1])
        elif name in ("swap", "SWAP"):
            c.swap(idx[0], idx[1])
        elif name in ("ccx", "CCX", "toffoli"):
            a, b, t = idx[0], idx[1], idx[2]
            for kind, x, y in (("H", t, 0), ("cx", b, t), ("Tdg", t, 0), ("cx", a, t), ("T", t, 0), ("cx", b, t),
                               ("Tdg", t, 0), ("cx", a, t), ("T", b, 0), ("T", t, 0), ("cx", a, b), ("H", t, 0),
                               ("T", a, 0), ("Tdg", b, 0), ("cx", a, b)):
                if kind == "cx":
                    c.cnot(x, y)
                else:
                    c.push(getattr(Gate, kind), x)
        else:
            raise self.err(f"unknown gate '{name}'")

    def _ref(self, ref: str, regs: Dict[str, Tuple[int, int]], what: str, reg_word: str) -> int:
        """resolve_qubit / resolve_cbit, src/qasm.rs:817-899 (`what` is
        "qubit" or "classical bit", `reg_word` "qubit" or "classical")."""
        if "[" in ref:
            name = ref[: ref.index("[")].strip()
            if "]" not in ref:
                raise self.err(f"invalid {what} reference '{ref}'")
            s = ref[ref.index("[") + 1: ref.index("]")]
            try:
                i = int(s.strip())
                if i < 0 or not s.strip().isdigit():
                    raise ValueError
            except ValueError:
                raise self.err(f"invalid {what} index '{s}'") from None
            if name not in regs:
                raise self.err(f"unknown {reg_word} register '{name}'")
            start, size = regs[name]
            if i >= size:
                raise self.err(f"{what} index {i} out of range for register '{name}' (size {size})")
            return start + i
        if ref not in regs:
            raise self.err(f"unknown {reg_word} register '{ref}'")
        return regs[ref][0]

    def qubit(self, ref: str) -> int:
        return self._ref(ref, self.qregs, "qubit", "qubit")

    def cbit(self, ref: str) -> int:
        return self._ref(ref, self.cregs, "classical bit", "classical")


def parse_qasm(source: str) -> Circuit:
    """parse_qasm, src/qasm.rs:73-76."""
    return _Parser(source).parse()


# ---------------------------------------------------------------------------
# Export (src/qasm_export.rs)
# ---------------------------------------------------------------------------
def format_angle(angle: float) -> str:
    """format_angle, src/qasm_export.rs:264-290: 0, +-pi/k, else 10 decimals
    with trailing zeros trimmed."""
    if abs(angle) < 1e-12:
        return "0"
    ratio = angle / math.pi
    if abs(ratio) > 1e-10:
        denom = int(_round_half_away(1.0 / ratio))
        if denom != 0 and abs(ratio - 1.0 / denom) < 1e-10:
            if denom == 1:
                return "pi"
            if denom == -1:
                return "-pi"
            return f"pi/{denom}" if denom > 0 else f"-pi/{-denom}"
    return f"{angle:.10f}".rstrip("0").rstrip(".")


def _round_half_away(x: float) -> float:
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def _gate_qasm(gid: int, theta: float) -> str:
    name = _GATE_NAMES[gid]
    plain = {"I": "id", "X": "x", "Y": "y", "Z": "z", "H": "h", "S": "s", "Sdg": "sdg", "T": "t", "Tdg": "tdg"}
    if name in plain:
        return plain[name]
    return f"{name.lower()}({format_angle(theta)})"


def _analyze(c: Circuit) -> Tuple[int, int, bool]:
    max_q = max_c = 0
    meas = False
    for k, g, a, b, t in c._ops:
        if k in (1, 2, 3, 7):
            max_q = max(max_q, a, b)
        elif k == 5:
            max_q, max_c, meas = max(max_q, a), max(max_c, b), True
        elif k == 6:
            max_q, max_c = max(max_q, a), max(max_c, b & 0xFFFF)
        else:
            max_q = max(max_q, a)
    return (max_q + 1 if c._ops else 0), max_c + 1, meas


def _export(c: Circuit, nq: int) -> str:
    _, ncb, meas = _analyze(c)
    lines = ["OPENQASM 2.0;", 'include "qelib1.inc";', ""]
    if nq > 0:
        lines.append(f"qreg q[{nq}];")
    if meas and ncb > 0:
        lines.append(f"creg c[{ncb}];")
    lines.append("")
    for k, g, a, b, t in c._ops:
        if k == 0:
            lines.append(f"{_gate_qasm(g, t)} q[{a}];")
        elif k == 1:
            lines.append(f"cx q[{a}], q[{b}];")
        elif k == 2:
            lines.append(f"cz q[{a}], q[{b}];")
        elif k == 3:
            lines.append(f"cx q[{a}], q[{b}];\ncx q[{b}], q[{a}];\ncx q[{a}], q[{b}];")
        elif k == 5 and meas:
            lines.append(f"measure q[{a}] -> c[{b}];")
        elif k == 6 and meas:
            lines.append(f"if (c[{b & 0xFFFF}] == {(b >> 16) & 1}) {_gate_qasm(g, t)} q[{a}];")
        elif k == 7:
            raise ValueError("the exact controlled phase has no QASM 2.0 form in the crate's exporter")
    return "\n".join(lines)


def export_qasm(circuit: Circuit) -> str:
    """export_qasm, src/qasm_export.rs:31-34 (channels are dropped; SWAP is
    written as three CNOTs)."""
    return _export(circuit, _analyze(circuit)[0])


def export_qasm_with_qubits(circuit: Circuit, num_qubits: int) -> str:
    """export_qasm_with_qubits, src/qasm_export.rs:42-56."""
    if _analyze(circuit)[0] > num_qubits:
        raise ValueError("circuit uses a qubit outside the declared register")
    return _export(circuit, num_qubits)


# ---------------------------------------------------------------------------
# JSON SerializedCircuit (src/serialize.rs)
# ---------------------------------------------------------------------------
def _gate_json(gid: int, theta: float) -> Tuple[str, list]:
    name = _GATE_NAMES[gid]
    return name, ([float(theta)] if name in ("Rx", "Ry", "Rz") else [])


def _serialized(c: Circuit) -> dict:
    max_q = max_c = 0
    ops = []
    for i, (k, g, a, b, t) in enumerate(c._ops):
        if k == 0:
            name, params = _gate_json(g, t)
            ops.append({"type": "single", "gate": name, "target": a, "params": params})
            max_q = max(max_q, a)
        elif k == 1:
            ops.append({"type": "cx", "control": a, "target": b})
            max_q = max(max_q, a, b)
        elif k == 2:
            ops.append({"type": "cz", "first": a, "second": b})
            max_q = max(max_q, a, b)
        elif k == 3:
            ops.append({"type": "swap", "first": a, "second": b})
            max_q = max(max_q, a, b)
        elif k == 5:
            ops.append({"type": "measure", "target": a, "classical_bit": b})
            max_q, max_c = max(max_q, a), max(max_c, b)
        elif k == 6:
            name, params = _gate_json(g, t)
            ops.append({"type": "conditional", "classical_bit": b & 0xFFFF, "value": (b >> 16) & 1, "gate": name,
                        "target": a, "params": params})
            max_q, max_c = max(max_q, a), max(max_c, b & 0xFFFF)
        elif k == 4:
            ch = c._channels.get(i, NoiseChannel("Depolarizing", (t,)))
            ops.append({"type": "channel", "channel": ch.name, "target": a, "params": [float(x) for x in ch.params]})
            max_q = max(max_q, a)
        else:
            raise ValueError("the exact controlled phase has no SerializedOp form")
    return {"num_qubits": max_q + 1, "num_classical_bits": max_c + 1, "operations": ops}


def to_json(circuit: Circuit) -> str:
    """Circuit::to_json, src/serialize.rs:91-95 (pretty JSON)."""
    return json.dumps(_serialized(circuit), indent=2)


def to_json_bytes(circuit: Circuit) -> bytes:
    """Circuit::to_json_bytes, src/serialize.rs:97-101 (compact JSON)."""
    return json.dumps(_serialized(circuit), separators=(",", ":")).encode()


def _gate_from(name: str, params: list) -> Gate:
    if name in ("Rx", "Ry", "Rz"):
        return Gate(name, float(params[0]) if params else 0.0)
    if name in ("I", "X", "Y", "Z", "H", "S", "Sdg", "T", "Tdg"):
        return getattr(Gate, name)
    raise ValueError(f"Unknown gate: {name}")


def from_json(text: str) -> Circuit:
    """Circuit::from_json / SerializedCircuit::to_circuit, src/serialize.rs:155-226."""
    d = json.loads(text)
    c = Circuit()
    for op in d["operations"]:
        t = op["type"]
        if t == "single":
            c.push(_gate_from(op["gate"], op.get("params", [])), int(op["target"]))
        elif t == "cx":
            c.cnot(int(op["control"]), int(op["target"]))
        elif t == "cz":
            c.cz(int(op["first"]), int(op["second"]))
        elif t == "swap":
            c.swap(int(op["first"]), int(op["second"]))
        elif t == "measure":
            c.measure(int(op["target"]), int(op["classical_bit"]))
        elif t == "conditional":
            c.if_classical_bit(int(op["classical_bit"]), int(op["value"]),
                               _gate_from(op["gate"], op.get("params", [])), int(op["target"]))
        elif t == "channel":
            name = op["channel"]
            p = [float(x) for x in op.get("params", [])]
            if name not in NoiseChannel.KINDS:
                raise ValueError(f"Unknown noise channel: {name}")
            want = 2 if name == "ThermalDamping" else 1
            p = (p + [0.0] * want)[:want]
            c.channel(int(op["target"]), NoiseChannel(name, tuple(p)))
        else:
            raise ValueError(f"unknown variant `{t}`")
    return c


def from_json_bytes(data: bytes) -> Circuit:
    return from_json(data.decode())
