"""Yosys JSON netlist input (crates/nextpnr/src/frontend/mod.rs) and its
flattening into the placer's `FlatDesign`.

`parse_json` follows the reference parser step for step (mod.rs:172-238):
one net per signal bit of the top module (cells first, then the module's
ports), the constant nets and their driver cells, the cells with their
ports, parameters and attributes, one pseudo-cell per top-level port bit,
and last the names of the `netnames` section.  The reference reads JSON
objects into serde_json maps without `preserve_order` (`Cargo.toml`:
`serde_json = "1"`), so it walks every object in sorted key order; the same
order is used here, which fixes the order in which cells and nets are
created and which module is "first".

`to_flat_design` is this package's own bridge from a parsed netlist to the
placer: a Yosys netlist carries no device geometry, so the grid, the IO
locations and the initial positions are chosen here (see its docstring).
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

IN, OUT, INOUT = "in", "out", "inout"  # PortType
ZERO, ONE, UNDEF = "0", "1", None      # BitValue constants; a signal is its int index

GND_NET, VCC_NET = "$PACKER_GND_NET", "$PACKER_VCC_NET"
GND_CELL, VCC_CELL = "$PACKER_GND", "$PACKER_VCC"
IO_TYPES = ("$nextpnr_IBUF", "$nextpnr_OBUF", "$nextpnr_IOBUF")


class NetlistError(ValueError):
    """A malformed or unsupported Yosys JSON netlist (the reference's `bail!`)."""


def _is_int(v) -> bool:
    return isinstance(v, int) and not isinstance(v, bool)


def _i64(v, what: str) -> int:
    # serde_json's as_i64: JSON integers in the i64 range only
    if not _is_int(v) or not (-(1 << 63) <= v < (1 << 63)):
        raise NetlistError(f"{what} is not a valid integer: {v!r}")
    return v


def parse_bit_value(v):
    """mod.rs:30-45: an int is a signal index; "0", "1" are constants; "x"
    and "z" are unconnected (None)."""
    if _is_int(v) or isinstance(v, float):
        return _i64(v, "Bit index")
    if isinstance(v, str):
        if v == "0":
            return ZERO
        if v == "1":
            return ONE
        if v in ("x", "z"):
            return UNDEF
        raise NetlistError(f"constant bit {v!r} is not one of 0, 1, x, z")
    raise NetlistError(f"a connection bit must be a net index or a constant string, got {v!r}")


def infer_port_direction(cell_type: str, port: str) -> str:
    """mod.rs:49-71: direction of a port whose cell lacks `port_directions`."""
    if cell_type == "LUT4":
        return OUT if port == "F" else IN
    if cell_type == "CARRY4":
        return OUT if port.startswith("CO") or port.startswith("O") else IN
    if cell_type in ("GND", "VCC", "GND_DRV", "VCC_DRV"):
        return OUT
    if cell_type == "IOB":
        return INOUT
    return OUT if port in ("Q", "O", "F", "Y") else IN


def parse_port_direction(d: str) -> str:
    """mod.rs:74-81."""
    try:
        return {"input": IN, "output": OUT, "inout": INOUT}[d]
    except (KeyError, TypeError):
        raise NetlistError(f"port direction {d!r} is not input, output or inout") from None


@dataclass(frozen=True)
class Prop:
    """A parameter or attribute value (netlist/property.rs): kind "int",
    "bits" (a string of 0/1) or "str"."""

    kind: str
    value: object

    def is_string(self) -> bool:
        return self.kind == "str"

    def as_int(self) -> Optional[int]:
        """property.rs:38-52: ints as they are, strings by decimal parse,
        bit vectors as binary (0 when empty or out of i64 range)."""
        if self.kind == "int":
            return int(self.value)
        if self.kind == "str":
            s = str(self.value)
            t = s[1:] if s[:1] in "+-" else s
            if not t or not t.isdigit() or not t.isascii():
                return None
            v = int(s)
            return v if -(1 << 63) <= v < (1 << 63) else None
        bits = str(self.value)
        if not bits:
            return 0
        if all(c in "01" for c in bits):
            v = int(bits, 2)
            return v if v < (1 << 63) else 0
        return None

    def as_str(self) -> str:
        return str(self.value)


def parse_property(v) -> Prop:
    """mod.rs:94-112: JSON ints are ints; a non-empty string of only 0 and 1
    is a bit vector; any other string is a string."""
    if _is_int(v) or isinstance(v, float):
        return Prop("int", _i64(v, "Property number"))
    if isinstance(v, str):
        if v and all(c in "01" for c in v):
            return Prop("bits", v)
        return Prop("str", v)
    raise NetlistError(f"property values must be strings or integers, got {v!r}")


def port_bit_name(base: str, i: int, total: int) -> str:
    """mod.rs:149-155: `port` for a one-bit port, else `port[i]`."""
    return base if total == 1 else f"{base}[{i}]"


@dataclass
class Port:
    name: str
    type: str
    net: Optional[int] = None    # net index
    user: Optional[int] = None   # index among the net's users (inputs and inouts)


@dataclass
class Cell:
    name: str
    type: str
    ports: Dict[str, Port] = field(default_factory=dict)
    params: Dict[str, Prop] = field(default_factory=dict)
    attrs: Dict[str, Prop] = field(default_factory=dict)

    def port_net(self, port: str) -> Optional[int]:
        p = self.ports.get(port)
        return None if p is None else p.net

    def port_type(self, port: str) -> Optional[str]:
        p = self.ports.get(port)
        return None if p is None else p.type


@dataclass
class Net:
    name: str
    driver: Optional[Tuple[int, str]] = None            # (cell index, port)
    users: List[Tuple[int, str]] = field(default_factory=list)
    attrs: Dict[str, Prop] = field(default_factory=dict)


@dataclass
class Netlist:
    """The top module as the reference's `Design` holds it: cells and nets
    in creation order, looked up by name."""

    top_module: str = ""
    cells: List[Cell] = field(default_factory=list)
    nets: List[Net] = field(default_factory=list)
    _cell_ix: Dict[str, int] = field(default_factory=dict)
    _net_ix: Dict[str, int] = field(default_factory=dict)

    def add_cell(self, name: str, cell_type: str) -> int:
        if name in self._cell_ix:
            raise NetlistError(f"cell already exists in design: {name!r}")
        self._cell_ix[name] = len(self.cells)
        self.cells.append(Cell(name, cell_type))
        return len(self.cells) - 1

    def add_net(self, name: str) -> int:
        if name in self._net_ix:
            raise NetlistError(f"net already exists in design: {name!r}")
        self._net_ix[name] = len(self.nets)
        self.nets.append(Net(name))
        return len(self.nets) - 1

    def rename_net(self, ni: int, name: str) -> None:
        """design.rs:240-256: a name taken by another net is an error."""
        old = self.nets[ni].name
        if old == name:
            return
        other = self._net_ix.get(name)
        if other is not None and other != ni:
            raise NetlistError(f"cannot rename net to an already-used name: {name!r}")
        del self._net_ix[old]
        self._net_ix[name] = ni
        self.nets[ni].name = name

    def cell_by_name(self, name: str) -> Optional[int]:
        return self._cell_ix.get(name)

    def net_by_name(self, name: str) -> Optional[int]:
        return self._net_ix.get(name)

    def cell(self, ci: int) -> Cell:
        return self.cells[ci]

    def net(self, ni: int) -> Net:
        return self.nets[ni]

    @property
    def num_cells(self) -> int:
        return len(self.cells)

    @property
    def num_nets(self) -> int:
        return len(self.nets)

    def connect(self, ci: int, port: str, ptype: str, ni: int) -> None:
        """mod.rs:424-449: an output drives the net; an input or inout is
        appended to its users."""
        p = self.cells[ci].ports[port]
        if ptype == OUT:
            p.net, p.user = ni, None
            self.nets[ni].driver = (ci, port)
        else:
            p.net, p.user = ni, len(self.nets[ni].users)
            self.nets[ni].users.append((ci, port))


def _obj(v, what: str) -> dict:
    if not isinstance(v, dict):
        raise NetlistError(f"{what} must be an object")
    return v


def _sorted(d: dict):
    return sorted(d.items(), key=lambda kv: kv[0])


def _find_top(modules: dict) -> Tuple[str, dict]:
    """mod.rs:119-143: the first module (in key order) whose `top`
    attribute is a nonzero integer, else the first module."""
    if not modules:
        raise NetlistError("the netlist defines no module")
    for name, mod in _sorted(modules):
        attrs = mod.get("attributes") if isinstance(mod, dict) else None
        if isinstance(attrs, dict) and "top" in attrs:
            v = parse_property(attrs["top"]).as_int()
            if v is not None and v != 0:
                return name, mod
    name, mod = _sorted(modules)[0]
    return name, mod


def _net_of(bit, bit_to_net: Dict[int, int], gnd: int, vcc: int) -> Optional[int]:
    if bit is ZERO:
        return gnd
    if bit is ONE:
        return vcc
    if bit is UNDEF:
        return None
    if bit not in bit_to_net:
        raise NetlistError("a connection names a bit that no net carries")
    return bit_to_net[bit]


def parse_json(text: str) -> Netlist:
    """mod.rs:172-238: the top module of a Yosys `write_json` netlist."""
    try:
        root = json.loads(text)
    except (ValueError, TypeError) as e:
        raise NetlistError(f"the netlist is not valid JSON: {e}") from e
    if not isinstance(root, dict) or not isinstance(root.get("modules"), dict):
        raise NetlistError("the netlist has no 'modules' object")
    top_name, mod = _find_top(root["modules"])
    nl = Netlist(top_module=top_name)
    mod = mod if isinstance(mod, dict) else {}
    cells = mod.get("cells") if isinstance(mod.get("cells"), dict) else {}
    ports = mod.get("ports") if isinstance(mod.get("ports"), dict) else {}

    # step 1 (mod.rs:241-286): one net per signal bit, cells' connections first
    bit_to_net: Dict[int, int] = {}

    def see(bits, what):
        if not isinstance(bits, list):
            raise NetlistError(f"the bits of {what} are not a list")
        for b in bits:
            v = parse_bit_value(b)
            if _is_int(v) and v not in bit_to_net:
                bit_to_net[v] = nl.add_net(f"$signal${v}")

    for _, cj in _sorted(cells):
        conns = cj.get("connections") if isinstance(cj, dict) else None
        if isinstance(conns, dict):
            for _, bits in _sorted(conns):
                see(bits, "Cell connection")
    for _, pj in _sorted(ports):
        if isinstance(pj, dict) and isinstance(pj.get("bits"), list):
            see(pj["bits"], "Port")

    # step 2 (mod.rs:289-316): constant nets and their driver cells
    gnd = nl.add_net(GND_NET)
    vcc = nl.add_net(VCC_NET)
    for cname, ctype, ni in ((GND_CELL, "GND", gnd), (VCC_CELL, "VCC", vcc)):
        ci = nl.add_cell(cname, ctype)
        nl.cells[ci].ports["Y"] = Port("Y", OUT)
        nl.connect(ci, "Y", OUT, ni)

    # step 3 (mod.rs:319-419): cells
    for cname, cj in _sorted(cells):
        cj = _obj(cj, f"cell {cname!r}")
        ctype = cj.get("type")
        if not isinstance(ctype, str):
            raise NetlistError(f"cell {cname!r} has no 'type'")
        ci = nl.add_cell(cname, ctype)
        dirs = {}
        if isinstance(cj.get("port_directions"), dict):
            for pname, dv in cj["port_directions"].items():
                if not isinstance(dv, str):
                    raise NetlistError("a cell's port direction is not a string")
                dirs[pname] = parse_port_direction(dv)
        if isinstance(cj.get("connections"), dict):
            for pname, bits in _sorted(cj["connections"]):
                if not isinstance(bits, list):
                    raise NetlistError("a cell connection's bits are not a list")
                ptype = dirs.get(pname) or infer_port_direction(ctype, pname)
                for i, b in enumerate(bits):
                    bit = parse_bit_value(b)
                    pn = port_bit_name(pname, i, len(bits))
                    nl.cells[ci].ports.setdefault(pn, Port(pn, ptype))
                    ni = _net_of(bit, bit_to_net, gnd, vcc)
                    if ni is not None:
                        nl.connect(ci, pn, ptype, ni)
        for key, dst in (("parameters", nl.cells[ci].params), ("attributes", nl.cells[ci].attrs)):
            if isinstance(cj.get(key), dict):
                for k, v in _sorted(cj[key]):
                    dst[k] = parse_property(v)

    # step 4 (mod.rs:451-525): a pseudo-cell per top-level port bit
    io = {IN: ("$nextpnr_IBUF", "O", OUT), OUT: ("$nextpnr_OBUF", "I", IN),
          INOUT: ("$nextpnr_IOBUF", "IO", INOUT)}
    for pname, pj in _sorted(ports):
        pj = _obj(pj, f"port {pname!r}")
        if not isinstance(pj.get("direction"), str):
            raise NetlistError(f"top-level port {pname!r} has no 'direction'")
        pdir = parse_port_direction(pj["direction"])
        bits = pj.get("bits")
        if not isinstance(bits, list):
            raise NetlistError(f"top-level port {pname!r} has no 'bits'")
        ctype, iport, itype = io[pdir]
        for i, b in enumerate(bits):
            bit = parse_bit_value(b)
            ci = nl.add_cell(f"$io${port_bit_name(pname, i, len(bits))}", ctype)
            nl.cells[ci].ports[iport] = Port(iport, itype)
            ni = _net_of(bit, bit_to_net, gnd, vcc)
            if ni is not None:
                nl.connect(ci, iport, itype, ni)

    # step 5 (mod.rs:532-585): names and attributes of `netnames`
    names = mod.get("netnames") if isinstance(mod.get("netnames"), dict) else {}
    for nname, nj in _sorted(names):
        nj = _obj(nj, f"netname {nname!r}")
        hide = nj.get("hide_name", 0)
        if (hide if _is_int(hide) else 0) != 0:
            continue
        bits = nj.get("bits")
        if not isinstance(bits, list):
            raise NetlistError("a netnames entry has no 'bits'")
        for i, b in enumerate(bits):
            v = parse_bit_value(b)
            if _is_int(v) and v in bit_to_net:
                nl.rename_net(bit_to_net[v], port_bit_name(nname, i, len(bits)))
        if isinstance(nj.get("attributes"), dict) and bits:
            v = parse_bit_value(bits[0])
            if _is_int(v) and v in bit_to_net:
                for k, pv in _sorted(nj["attributes"]):
                    nl.nets[bit_to_net[v]].attrs[k] = parse_property(pv)
    return nl


def _ring(W: int, H: int) -> List[Tuple[int, int]]:
    """The grid's perimeter tiles: along the bottom, up the right side, back
    along the top and down the left side."""
    ring = [(x, 0) for x in range(W)]
    ring += [(W - 1, y) for y in range(1, H)]
    if H > 1:
        ring += [(x, H - 1) for x in range(W - 2, -1, -1)]
    if W > 1:
        ring += [(0, y) for y in range(H - 2, 0, -1)]
    return ring


def to_flat_design(nl: Netlist, W: Optional[int] = None, seed: int = 0, name: str = ""):
    """A placement problem for the parsed netlist.

    A Yosys netlist has no device, so this bridge chooses the geometry the
    way this package's synthetic designs do:
    - movable cells: every cell but the top-level port pseudo-cells and the
      constant drivers, each of unit area;
    - fixed cells: the port pseudo-cells, spread evenly along the perimeter
      of the W x W tile grid, in creation order;
    - nets: every net with at least two pins on those cells (the constant
      nets are left out, as a packer folds constants into the cells that use
      them); a pin per driver and per user;
    - W: ceil(sqrt(movable / 0.7)) (at least 4) unless given, bins the
      nearest power of two;
    - initial positions: the movable cells on distinct tiles, shuffled with
      `seed`.
    Returns (FlatDesign, names of the movable cells, names of the fixed cells).
    """
    from .placer import FlatDesign, bins_for_side

    mov, fix = [], []
    for ci, c in enumerate(nl.cells):
        if c.type in IO_TYPES:
            fix.append(ci)
        elif c.name not in (GND_CELL, VCC_CELL):
            mov.append(ci)
    if W is None:
        W = max(4, int(math.ceil(math.sqrt(len(mov) / 0.7)))) if mov else 4
    if len(mov) > W * W:
        raise NetlistError(f"{len(mov)} movable cells do not fit a {W} x {W} grid")
    slot = {ci: k for k, ci in enumerate(mov)}
    slot.update({ci: len(mov) + k for k, ci in enumerate(fix)})
    net_off, pin_cell = [0], []
    for n in nl.nets:
        if n.name in (GND_NET, VCC_NET):
            continue
        pins = ([n.driver] if n.driver is not None else []) + list(n.users)
        cells = [slot[ci] for ci, _ in pins if ci in slot]
        if len(cells) < 2:
            continue
        pin_cell += cells
        net_off.append(len(pin_cell))
    ring = _ring(W, W)
    fx = np.array([ring[int(k * len(ring) / len(fix))][0] for k in range(len(fix))], np.float64)
    fy = np.array([ring[int(k * len(ring) / len(fix))][1] for k in range(len(fix))], np.float64)
    tiles = np.random.default_rng(seed).permutation(W * W)[:len(mov)]
    d = FlatDesign(W=W, H=W, Bx=bins_for_side(W), By=bins_for_side(W),
                   x0=(tiles % W).astype(np.float64), y0=(tiles // W).astype(np.float64),
                   fixed_x=fx, fixed_y=fy, net_off=np.array(net_off, np.uint32),
                   pin_cell=np.array(pin_cell, np.uint32), name=name or nl.top_module)
    return d, [nl.cells[ci].name for ci in mov], [nl.cells[ci].name for ci in fix]


def read_json(path: str) -> Netlist:
    """parse_json of a file."""
    with open(path, "r", encoding="utf-8") as f:
        return parse_json(f.read())
