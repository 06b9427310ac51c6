"""Config front end (CPU): the Lua-subset reader of config.lua-style files
and the catalogue of initial conditions by the reference's names."""
import os

import numpy as np
import pytest

from hydrodynamicsgpu_amd import config, ics, initconds, luaconfig


def _g(text):
    return luaconfig.evaluate(text, initconds.lua_setups())


def test_comment_forms_and_block_toggles():
    g = _g("""
--[[ 3D
size = {32, 32, 32}
--]]
-- [[ 2D
size = {256, 256}
--]]
--[=[ nested
--[==[ inner ]==]
x = 5
--]=]
y = 'a' -- trailing comment
""")
    assert luaconfig.to_py(g.get("size")) == [256.0, 256.0]
    assert g.get("x") is None
    assert g.get("y") == "a"


def test_expressions_follow_lua():
    g = _g("""
a = 1 + 2 * 3 ^ 2
b = nil or 5
c = false and 1
d = #{1, 2, 3}
e = 'x' .. 1 .. 2.5
f = math.sqrt(16) + math.pi * 0
h = 7/5
i = not nil
j = 2 ^ 3 ^ 2
k = -2 ^ 2
t = {1, 2, n = 3, [4] = 'four'}
t[3] = 'three'
m = #t
""")
    assert g.get("a") == 19
    assert g.get("b") == 5
    assert g.get("c") is False
    assert g.get("d") == 3
    assert g.get("e") == "x12.5"
    assert g.get("f") == 4
    assert g.get("h") == 7 / 5
    assert g.get("i") is True
    assert g.get("j") == 512  # right associative
    assert g.get("k") == -4   # unary minus binds looser than ^
    assert g.get("m") == 4
    assert g.get("t").get("n") == 3


def test_if_chains_locals_and_field_assignment():
    g = _g("""
size = {64, 48}
local k = #size
size[k] = size[k] * 3
if #size == 1 then mode = 'one'
elseif #size == 2 then mode = 'two'
else mode = 'three' end
xmid = {}
for_me = k
xmid[1] = .5 * (-1 + 1)
""")
    assert luaconfig.to_py(g.get("size")) == [64, 144]
    assert g.get("mode") == "two"
    assert g.get("k") is None  # locals do not leak
    assert g.get("for_me") == 2


_CONFIG = """
package.path = '?.lua;?/?.lua'
require 'util'
require 'initConds'
slopeLimiterName = 'MinMod'   -- the reference reads `slopeLimiter`, not this
integratorName = 'RungeKutta4'
xmin = {-.5, -.5, -.5}
xmax = {.5, .5, .5}
useFixedDT = false
fixedDT = .001
boundaryMethods = {
    {min='MIRROR', max='MIRROR'},
    {min='FREEFLOW', max='PERIODIC'},
    {min='FREEFLOW', max='FREEFLOW'},
}
defs = {
    idealGas_heatCapacityRatio = 1.4,
    srhd_solvePrimMaxIter = 500,
}
--[[ 3D
size = {32, 32, 32}
--]]
-- [[ 2D
size = {64, 48}
--]]
solverName = 'EulerHLLC'
initCondName = 'Configuration 3'
initConds[initCondName].setup()
if solverName == 'SRHDRoe' then cfl = .1 end
cfl = cfl or 1/#size
"""


def test_load_config_text_maps_every_key():
    d, info = luaconfig.load(_CONFIG)
    assert d.dim == 2 and list(d.size)[:2] == [64, 48]
    assert d.solver == config.SOLVERS["HLLC"] and d.equation == config.EQUATIONS["Euler"]
    assert d.integrator == config.INTEGRATORS.index("RungeKutta4")
    # slopeLimiterName is not the key the reference reads: default Superbee
    assert d.limiter == config.LIMITERS.index("Superbee")
    assert info["slopeLimiterName"] == "MinMod"
    # the configuration's setup() set cfl and free-flow boundaries
    assert d.cfl == pytest.approx(.475)
    assert all(d.boundary[a][m] == config.BOUNDARY["FREEFLOW"] for a in range(3) for m in range(2))
    assert d.srhd_solve_prim_max_iter == 500
    assert info["initCondName"] == "Configuration 3"
    q = info["initState"](d).reshape(48, 64, 4)
    # quadrant q1 (x > 0, y > 0): rho 1.5, p 1.5, at rest
    assert q[40, 50, 0] == 1.5 and q[40, 50, 1] == 0 and q[40, 50, 3] == pytest.approx(1.5 / 0.4)
    # q3 (x < 0, y < 0): rho .138, v = (1.206, 1.206)
    assert q[5, 5, 0] == .138 and q[5, 5, 1] == pytest.approx(.138 * 1.206)


def test_app_default_cfl_without_config_line():
    d, _ = luaconfig.load("size = {32}\nsolverName = 'EulerRoe'\n")
    assert d.cfl == pytest.approx(0.5)  # HydroGPUApp's own default
    assert d.dim == 1


def test_unsupported_configs_are_reported():
    with pytest.raises(luaconfig.LuaError, match="calcSolid"):
        luaconfig.load("size = {32, 32}\nfunction calcSolid(x,y,z) return 0 end\n")
    with pytest.raises(ValueError, match="ADM"):
        luaconfig.load("size = {32, 32}\ninitConds['NR Gauge Shock Waves'].setup()\n")


REF_CONFIG = "/root/reference/res/config.lua"


@pytest.mark.skipif(not os.path.exists(REF_CONFIG), reason="reference sources not present")
def test_reference_config_file():
    # the reference's shipped configuration: 2D EulerRoe Sod, ForwardEuler
    d, info = luaconfig.load(REF_CONFIG)
    assert info["initCondName"] == "Sod"
    assert d.dim == 2 and list(d.size)[:2] == [256, 256]
    assert d.solver == config.SOLVERS["Roe"] and d.equation == config.EQUATIONS["Euler"]
    assert d.integrator == config.INTEGRATORS.index("ForwardEuler")
    assert d.limiter == config.LIMITERS.index("Superbee")
    assert d.cfl == pytest.approx(0.5)  # cfl = cfl or 1/#size
    assert d.gamma == 1.4
    assert not info["useGravity"]


# ---------------------------------------------------------------- catalogue
def test_catalogue_sod_matches_the_bench_generator():
    d = config.make_desc([1024], xmin=[-1], xmax=[1])
    cfg, fn = initconds.setup("Sod", {"size": [1024]})
    assert cfg["defs"]["idealGas_heatCapacityRatio"] == 7 / 5
    np.testing.assert_array_equal(fn(d), ics.sod(d))


def test_catalogue_setups_change_the_globals():
    cfg, _ = initconds.setup("Double Mach Reflection", {"size": [64, 32]})
    assert cfg["size"] == [128, 16] and cfg["xmax"] == [4.0, 1.0, 1.0]
    assert cfg["boundaryMethods"][1] == {"min": "MIRROR", "max": "FREEFLOW"}
    cfg, _ = initconds.setup("Shock Bubble Interaction", {"size": [64, 32]})
    assert cfg["size"] == [128, 32] and cfg["xmin"] == [-1.0, -.5, -.5]
    cfg, _ = initconds.setup("Kelvin-Helmholtz", {"size": [64, 64]})
    assert all(b == {"min": "PERIODIC", "max": "PERIODIC"} for b in cfg["boundaryMethods"])
    cfg, _ = initconds.setup("Square Cavity", {"size": [64, 64]})
    assert cfg["boundaryMethods"][1] == {"min": "MIRROR", "max": "NONE"}
    with pytest.raises(ValueError, match="ADM"):
        initconds.setup("NR Stellar", {"size": [64, 64]})


def test_catalogue_solid_potential_and_gravity_channels():
    # Flow Around Cylinder: NONE on x min, solid disk r < .25 (res/initConds.lua:443-466)
    cfg, _ = initconds.setup("Flow Around Cylinder", {"size": [64, 64]})
    assert cfg["boundaryMethods"][0] == {"min": "NONE", "max": "FREEFLOW"}
    d = config.make_desc(cfg["size"], boundaryMethods=cfg["boundaryMethods"])
    q, pot, solid = initconds.channels_function("Flow Around Cylinder")(d)
    x, y, _ = ics.cell_centers(d)[:3]
    X, Y = np.meshgrid(x, y)
    np.testing.assert_array_equal(solid.reshape(64, 64), (X * X + Y * Y < .0625).astype(np.uint8))
    assert pot is None and np.allclose(q.reshape(-1, 4)[:, 1], .2)
    # Forward Facing Step: solid where x > xmid and y < ymid
    d = config.make_desc([40, 20])
    _, _, solid = initconds.channels_function("Forward Facing Step")(d)
    x, y, _ = ics.cell_centers(d)[:3]
    X, Y = np.meshgrid(x, y)
    np.testing.assert_array_equal(solid.reshape(20, 40), ((X > 0) & (Y < 0)).astype(np.uint8))
    # Rayleigh-Taylor: the last axis tripled (extent and cells), mirror there,
    # potential y - ymin, pressure 2.5 - rho phi, density 2 above the middle
    cfg, _ = initconds.setup("Rayleigh-Taylor", {"size": [32, 20]})
    assert cfg["size"] == [32, 60] and cfg["xmin"][1] == -1.5 and cfg["xmax"][1] == 1.5
    assert cfg["boundaryMethods"][1] == {"min": "MIRROR", "max": "MIRROR"}
    d = config.make_desc(cfg["size"], xmin=cfg["xmin"], xmax=cfg["xmax"])
    q, pot, solid = initconds.channels_function("Rayleigh-Taylor")(d)
    y = ics.cell_centers(d)[1]
    Y = np.repeat(y, 32)
    np.testing.assert_allclose(pot, Y + 1.5, rtol=0, atol=1e-15)
    rho = q.reshape(-1, 4)[:, 0]
    np.testing.assert_array_equal(rho, np.where(Y > 0, 2.0, 1.0))
    v = q.reshape(-1, 4)[:, 1:3] / rho[:, None]
    assert 0 < np.abs(v).max() <= .001 and solid is None
    eint = (q.reshape(-1, 4)[:, 3] - .5 * rho * (v * v).sum(1)) / rho
    np.testing.assert_allclose((d.gamma - 1) * rho * eint, 2.5 - rho * (Y + 1.5), rtol=1e-12)
    # self-gravitation tests: useGravity, density 1 inside the sources, .1 outside
    for name in ("self-gravitation test 1", "self-gravitation test 1 spinning", "self-gravitation test 2",
                 "self-gravitation test 4"):
        cfg, fn = initconds.setup(name, {"size": [40, 40]})
        assert cfg["useGravity"] is True
        d = config.make_desc([40, 40], useGravity=True)
        q = fn(d).reshape(-1, 4)
        assert set(np.unique(q[:, 0])) == {0.1, 1.0}, name
    d = config.make_desc([40, 40])
    q = initconds.state_function("self-gravitation test 2")(d).reshape(-1, 4)
    x, y, _ = ics.cell_centers(d)[:3]
    X, Y = np.meshgrid(x, y)
    np.testing.assert_allclose(q[:, 1], (-5 * q[:, 0] * Y.ravel()), rtol=1e-15)
    np.testing.assert_allclose(q[:, 2], (5 * q[:, 0] * X.ravel()), rtol=1e-15)
    # the Lua front end passes useGravity on
    desc, info = luaconfig.load("size = {32, 32}\ninitConds['self-gravitation test 4'].setup()\n")
    assert desc.use_gravity == 1 and info["initChannels"] is not None


def test_catalogue_energies():
    # Brio-Wu: total energy carries |B|^2 / 2 (buildStateEuler), gamma 2
    cfg, fn = initconds.setup("Brio-Wu", {"size": [64]})
    d = config.make_desc([64], defs=cfg["defs"])
    q = fn(d).reshape(64, 3)
    eint = 1.0 / ((2 - 1) * 1.0)
    assert q[0, 2] == pytest.approx(1.0 * eint + 0.5 * (.75 ** 2 + 1.0))
    # Marti & Muller #1 gives the specific internal energy, not the pressure
    cfg, fn = initconds.setup("Marti & Muller 2008 Problem #1", {"size": [64]})
    d = config.make_desc([64], defs=cfg["defs"], solverName="SRHDRoe")
    q = fn(d).reshape(64, 3)
    assert q[0, 0] == 10 and q[0, 2] == pytest.approx(10 * 2)
    assert q[-1, 0] == 1 and q[-1, 2] == pytest.approx(1e-6)
    # Constant: unit velocity and field in every used axis
    d = config.make_desc([16, 16])
    q = initconds.state_function("Constant")(d).reshape(-1, 4)
    assert np.allclose(q[:, 3], 1.0 * (1.0 + 1.0 / (d.gamma - 1)) + 1.5)


def test_catalogue_kelvin_helmholtz_layout():
    d = config.make_desc([64, 64])
    q = initconds.state_function("Kelvin-Helmholtz")(d).reshape(64, 64, 4)
    yc = ics.cell_centers(d)[1]
    inside = (yc > -.25) & (yc < .25)
    assert np.all(q[inside, :, 0] == 2) and np.all(q[~inside, :, 0] == 1)
    noise = 64 * 2e-5
    vx = q[:, :, 1] / q[:, :, 0]
    assert np.all(np.abs(vx[inside] + .5) <= noise + 1e-15) and np.all(np.abs(vx[~inside] - .5) <= noise + 1e-15)


def test_catalogue_every_entry_builds_finite_states():
    for name, (eqs, _, _) in initconds.CATALOG.items():
        solver = "MaxwellRoe" if "Maxwell" in eqs else "EulerRoe"
        size = [40, 24] if name not in ("Maxwell-1",) else [32]
        cfg, fn = initconds.setup(name, {"size": size})
        d = config.make_desc(cfg["size"], xmin=cfg.get("xmin"), xmax=cfg.get("xmax"), defs=cfg.get("defs"),
                             solverName=solver)
        q = fn(d)
        assert q.shape == (config.num_states(d) * config.volume(d),), name
        assert np.all(np.isfinite(q)), name
        if solver == "EulerRoe":
            assert np.all(q.reshape(-1, config.num_states(d))[:, 0] > 0), name
