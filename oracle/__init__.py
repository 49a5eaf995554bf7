"""CPU oracle for the RCWA solve path -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy complex128 restatement of the reference
solver's algorithm (Kaeryv/Khepri, ``khepri/`` package).  Every function
cites the reference file:line it follows.

Who may use it: ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- and only as the reference results that
the GPU path is compared with, or as the timed CPU baseline.  The product package ``khepri_amd`` never imports it; the GPU
path has no CPU fallback.

Pinning status of each function (what checks it, besides retyping):
  expansion_indices / reciproc / g_vectors   closed forms (square & hex), tests/test_oracle.py
  compute_kplanar                            closed form
  convolution_matrix                         brute-force DFT, constant pixmap -> eps*I, 1-D grating sinc,
                                             golden vectors from the reference (tests/golden)
  free_space / uniform / half-space / build_scatmat / redheffer / stack / incident / poynting
                                             Fresnel thin-film known answer (reference test_complex_eps),
                                             1-D transfer matrix for multilayers, energy conservation,
                                             golden vectors
  structured_layer_modes                     eigen residual + golden vectors (eigenvalue multisets, layer S)
  fields (translate / fourier_fields / longitudinal / idft)
                                             uniform-slab analytic standing wave, golden field volumes
  joint_subspace / extended layer            explicit stride restatement vs index-formula, golden twisted R/T
  gen_bzi_grid / Drawing                     closed forms
"""
