"""Island-layer cases of the analytic.npz fixture (shared by the fixture
generator and the tests; pure data)."""

ANALYTIC_CASES = [
    # (pw, lattice, islands, eps_host, depth)
    ((5, 5), "square",
     [{"type": "disc", "params": [0.1, -0.05, 0.3], "epsilon": 4.0},
      {"type": "rectangle", "params": [-0.45, -0.4, -0.1, 0.2], "epsilon": 2.5 + 0.1j}], 1.5, 0.4),
    ((3, 5), "square",
     [{"type": "disc", "params": [0.0, 0.0, 0.25], "epsilon": 6.0}], 1.0, 0.3),
    ((5, 5), "hexagonal",
     [{"type": "disc", "params": [0.2, 0.1, 0.2], "epsilon": 3.0}], 2.0, 0.5),
]
# (wavelength, te, tm, theta_deg, phi_deg)
ANALYTIC_SOURCES = [(1.6, 1.0, 0.0, 0.0, 0.0), (1.9, 1.0, 1.0, 20.0, 35.0), (2.3, 0.0, 1.0, 7.0, 200.0)]
# every case is the stack [A, G, A] with G a uniform layer (eps 1.2, depth 0.25)
GAP = (1.2, 0.25)
