"""Yosys JSON netlist input (eisenjoch_amd/yosys.py) against the behaviour the
reference's frontend tests pin (crates/nextpnr/tests/frontend_tests.rs, on
the same blinky, multi-bit and constants netlists, kept as data under
tests/golden/), and the flattening into a placement problem.  CPU only."""
import json
from pathlib import Path

import numpy as np
import pytest

from eisenjoch_amd import yosys as Y

GOLD = Path(__file__).parent / "golden"


def load(name):
    return Y.parse_json((GOLD / f"yosys_{name}.json").read_text())


# ---- value parsing (frontend_tests.rs:13-91)

def test_bit_values():
    assert Y.parse_bit_value(42) == 42
    assert Y.parse_bit_value("0") is Y.ZERO
    assert Y.parse_bit_value("1") is Y.ONE
    assert Y.parse_bit_value("x") is Y.UNDEF and Y.parse_bit_value("z") is Y.UNDEF
    for bad in ("2", "abc", True, None, [1], {"a": 1}, 1.5):
        with pytest.raises(Y.NetlistError):
            Y.parse_bit_value(bad)


def test_port_directions():
    assert Y.parse_port_direction("input") == Y.IN
    assert Y.parse_port_direction("output") == Y.OUT
    assert Y.parse_port_direction("inout") == Y.INOUT
    with pytest.raises(Y.NetlistError):
        Y.parse_port_direction("sideways")


def test_inferred_port_directions():
    assert Y.infer_port_direction("LUT4", "F") == Y.OUT
    assert Y.infer_port_direction("LUT4", "A") == Y.IN
    assert Y.infer_port_direction("CARRY4", "CO3") == Y.OUT
    assert Y.infer_port_direction("CARRY4", "O2") == Y.OUT
    assert Y.infer_port_direction("CARRY4", "CI") == Y.IN
    assert Y.infer_port_direction("VCC_DRV", "X") == Y.OUT
    assert Y.infer_port_direction("IOB", "PAD") == Y.INOUT
    for p in ("Q", "O", "F", "Y"):
        assert Y.infer_port_direction("DFF", p) == Y.OUT
    assert Y.infer_port_direction("DFF", "D") == Y.IN


def test_properties():
    assert Y.parse_property(42).as_int() == 42
    p = Y.parse_property("0000000000001111")
    assert p.kind == "bits" and p.as_int() == 15
    s = Y.parse_property("blinky.v:5")
    assert s.is_string() and s.as_str() == "blinky.v:5" and s.as_int() is None
    assert Y.parse_property("").is_string()          # empty is a string, not a bit vector
    assert Y.parse_property("-12").as_int() == -12   # strings parse as decimal
    assert Y.parse_property("1" * 64).as_int() == 0  # out of i64: 0, as from_str_radix's unwrap_or
    with pytest.raises(Y.NetlistError):
        Y.parse_property(True)
    with pytest.raises(Y.NetlistError):
        Y.parse_property([1])


def test_port_bit_names():
    assert Y.port_bit_name("A", 0, 1) == "A"
    assert Y.port_bit_name("D", 2, 4) == "D[2]"


# ---- the blinky netlist (frontend_tests.rs:99-369)

def test_blinky_top_and_cells():
    nl = load("blinky")
    assert nl.top_module == "top"
    for c in ("lut0", "$PACKER_GND", "$PACKER_VCC", "$io$clk", "$io$led"):
        assert nl.cell_by_name(c) is not None
    assert nl.num_cells == 5


def test_blinky_lut():
    nl = load("blinky")
    c = nl.cell(nl.cell_by_name("lut0"))
    assert c.type == "LUT4"
    assert c.port_type("A") == Y.IN and c.port_net("A") is not None
    assert c.port_type("Z") == Y.OUT and c.port_net("Z") is not None
    assert c.params["INIT"].as_int() == 15
    assert c.attrs["src"].as_str() == "blinky.v:5.3-5.20"
    gnd = nl.net_by_name("$PACKER_GND_NET")
    assert c.port_net("B") == gnd and c.port_net("C") == gnd and c.port_net("D") == gnd


def test_blinky_nets_and_io():
    nl = load("blinky")
    for n in ("$PACKER_GND_NET", "$PACKER_VCC_NET", "clk", "led"):
        assert nl.net_by_name(n) is not None
    assert nl.net_by_name("$signal$2") is None  # renamed by netnames
    assert nl.cell(nl.cell_by_name("$io$clk")).type == "$nextpnr_IBUF"
    assert nl.cell(nl.cell_by_name("$io$led")).type == "$nextpnr_OBUF"
    clk = nl.net(nl.net_by_name("clk"))
    assert clk.driver == (nl.cell_by_name("$io$clk"), "O")
    assert (nl.cell_by_name("lut0"), "A") in clk.users
    led = nl.net(nl.net_by_name("led"))
    assert led.driver == (nl.cell_by_name("lut0"), "Z")
    assert led.users == [(nl.cell_by_name("$io$led"), "I")]


def test_creation_order_follows_sorted_keys():
    # serde_json maps without preserve_order iterate in key order: the
    # constant drivers first, then cells and port pseudo-cells by name
    nl = load("blinky")
    assert [c.name for c in nl.cells] == ["$PACKER_GND", "$PACKER_VCC", "lut0", "$io$clk", "$io$led"]
    assert [n.name for n in nl.nets] == ["clk", "led", "$PACKER_GND_NET", "$PACKER_VCC_NET"]


def test_gnd_net_users_and_driver():
    nl = load("blinky")
    gnd = nl.net(nl.net_by_name("$PACKER_GND_NET"))
    assert gnd.driver == (nl.cell_by_name("$PACKER_GND"), "Y")
    lut = nl.cell_by_name("lut0")
    assert gnd.users == [(lut, "B"), (lut, "C"), (lut, "D")]
    # each user knows its index among the net's users
    c = nl.cell(lut)
    assert [c.ports[p].user for p in ("B", "C", "D")] == [0, 1, 2]


# ---- multi-bit ports and constants (frontend_tests.rs:371-521)

def test_multibit_ports_io_and_nets():
    nl = load("multibit")
    c = nl.cell(nl.cell_by_name("buf0"))
    for i in range(4):
        assert c.port_type(f"I[{i}]") == Y.IN and c.port_net(f"I[{i}]") is not None
        assert c.port_type(f"O[{i}]") == Y.OUT and c.port_net(f"O[{i}]") is not None
        assert nl.cell(nl.cell_by_name(f"$io$data_in[{i}]")).type == "$nextpnr_IBUF"
        assert nl.cell(nl.cell_by_name(f"$io$data_out[{i}]")).type == "$nextpnr_OBUF"
        ibuf = nl.cell_by_name(f"$io$data_in[{i}]")
        assert nl.net(c.port_net(f"I[{i}]")).driver == (ibuf, "O")


def test_constant_one_and_x():
    nl = load("constants")
    a = nl.cell(nl.cell_by_name("cell_a"))
    assert a.port_net("I") == nl.net_by_name("$PACKER_VCC_NET")
    b = nl.cell(nl.cell_by_name("cell_b"))
    assert "I" in b.ports and b.port_net("I") is None


# ---- modules, errors, inout, netnames (frontend_tests.rs:523-880)

def module(name="m", **kw):
    m = {"attributes": {}, "parameter_default_values": {}, "ports": {}, "cells": {}, "netnames": {}}
    m.update(kw)
    return name, m


def doc(*mods):
    return json.dumps({"creator": "Yosys", "modules": dict(mods)})


def test_empty_module_has_only_the_constants():
    nl = Y.parse_json(doc(module("empty")))
    assert nl.num_cells == 2 and nl.num_nets == 2


@pytest.mark.parametrize("text", ['{ "creator": "Yosys" }', '{ "modules": {} }', "not json at all",
                                  '{ "modules": 3 }'])
def test_errors(text):
    with pytest.raises(Y.NetlistError):
        Y.parse_json(text)


def test_top_module_by_attribute_else_first_by_name():
    nl = Y.parse_json(doc(module("sub"), module("main", attributes={"top": "00000000000000000000000000000001"})))
    assert nl.top_module == "main"
    nl = Y.parse_json(doc(module("zeta"), module("alpha")))
    assert nl.top_module == "alpha"   # no top attribute: the first key in order
    nl = Y.parse_json(doc(module("b", attributes={"top": 0}), module("c", attributes={"top": 1})))
    assert nl.top_module == "c"


def test_inout_port_and_cell():
    nl = Y.parse_json(doc(module(ports={"pad": {"direction": "inout", "bits": [7]}},
                                 cells={"u": {"type": "IOB", "connections": {"PAD": [7]}}})))
    io = nl.cell(nl.cell_by_name("$io$pad"))
    assert io.type == "$nextpnr_IOBUF" and io.port_type("IO") == Y.INOUT
    n = nl.net(io.port_net("IO"))
    assert n.driver is None and len(n.users) == 2  # inout ports are users
    assert nl.cell(nl.cell_by_name("u")).port_type("PAD") == Y.INOUT  # inferred for IOB


def test_multiple_parameters():
    nl = Y.parse_json(doc(module(cells={"c": {"type": "X", "parameters": {
        "A": "0101", "B": 7, "C": "text", "D": ""}, "connections": {}}})))
    p = nl.cell(nl.cell_by_name("c")).params
    assert p["A"].as_int() == 5 and p["B"].as_int() == 7
    assert p["C"].is_string() and p["D"].is_string()


def test_netnames_rename_hidden_and_attributes():
    cells = {"c": {"type": "LUT4", "connections": {"A": [4, 5], "F": [6]}}}
    names = {"bus": {"hide_name": 0, "bits": [4, 5], "attributes": {"keep": "1"}},
             "$auto$h": {"hide_name": 1, "bits": [6]}}
    nl = Y.parse_json(doc(module(cells=cells, netnames=names)))
    assert nl.net_by_name("bus[0]") is not None and nl.net_by_name("bus[1]") is not None
    assert nl.net_by_name("$signal$6") is not None      # hidden names are not applied
    assert nl.net_by_name("$auto$h") is None
    assert nl.net(nl.net_by_name("bus[0]")).attrs["keep"].as_int() == 1  # first bit's net only
    assert "keep" not in nl.net(nl.net_by_name("bus[1]")).attrs


def test_rename_to_a_used_name_is_an_error():
    cells = {"c": {"type": "LUT4", "connections": {"A": [4], "B": [5], "F": [6]}}}
    names = {"n": {"hide_name": 0, "bits": [4]}, "o": {"hide_name": 0, "bits": [4]},
             "$signal$6": {"hide_name": 0, "bits": [5]}}
    with pytest.raises(Y.NetlistError):
        Y.parse_json(doc(module(cells=cells, netnames=names)))


def test_two_cells_share_a_net():
    cells = {"a": {"type": "LUT4", "connections": {"A": [2], "F": [9]}},
             "b": {"type": "LUT4", "connections": {"A": [9], "F": [3]}}}
    nl = Y.parse_json(doc(module(cells=cells)))
    n = nl.net(nl.net_by_name("$signal$9"))
    assert n.driver == (nl.cell_by_name("a"), "F")
    assert n.users == [(nl.cell_by_name("b"), "A")]


# ---- flattening into a placement problem

def test_flat_design_of_blinky():
    d, mov, fix = Y.to_flat_design(load("blinky"))
    assert mov == ["lut0"] and fix == ["$io$clk", "$io$led"]
    assert d.n_mov == 1 and d.n_fixed == 2 and d.W == 4
    # clk: IBUF (fixed 1) drives lut0 (cell 0); led: lut0 drives the OBUF (fixed 2)
    assert d.net_off.tolist() == [0, 2, 4]
    assert d.pin_cell.tolist() == [1, 0, 0, 2]
    assert 0 <= d.x0[0] < d.W and 0 <= d.y0[0] < d.H


def test_flat_design_of_a_generated_netlist_keeps_every_pin():
    # a synthetic netlist: 300 LUTs in a chain plus fanout nets, 8 inputs
    cells, bit = {}, 100
    for i in range(300):
        cells[f"l{i:03d}"] = {"type": "LUT4", "connections": {"A": [bit + i], "B": [2 + i % 8], "F": [bit + i + 1]}}
    ports = {f"in{k}": {"direction": "input", "bits": [2 + k]} for k in range(8)}
    ports["out"] = {"direction": "output", "bits": [bit + 300]}
    nl = Y.parse_json(doc(module(cells=cells, ports=ports)))
    d, mov, fix = Y.to_flat_design(nl, seed=3)
    assert d.n_mov == 300 and d.n_fixed == 9
    sizes = np.diff(d.net_off.astype(np.int64))
    assert sizes.min() >= 2
    # the chain nets have 2 pins (the first one has only its user: 1 pin, dropped);
    # each input net drives 300/8 LUTs plus its IBUF
    assert int(sizes.sum()) == 299 * 2 + 1 * 2 + 8 * 1 + 300
    tiles = (d.y0 * d.W + d.x0).astype(np.int64)
    assert np.unique(tiles).size == 300  # distinct tiles
    assert np.all((d.fixed_x == 0) | (d.fixed_x == d.W - 1) | (d.fixed_y == 0) | (d.fixed_y == d.H - 1))
