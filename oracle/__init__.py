"""CPU FP64 oracle for the specific_particle_tracer tracing path.

THIS PACKAGE IS TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker (or the timed CPU baseline). The product package ``bem_amd`` never
imports it; the product path runs on the HIP library or raises.

It is a NumPy/SciPy restatement of the reference algorithm, function by
function, each citing the reference file:line it follows. Third-party
arithmetic on the path is called the same way the reference calls it:
``scipy.special.ellipkm1`` / ``ellipe`` for the complete elliptic integrals
(reference ``bem/_elliptic.py:28-39``) and ``numpy.linalg.solve`` for the
dense collocation solve (``bem/axisymmetric.py:442``).

Pinning status of each function (how a plausible mistake would be caught):

=====================================  =============================================
function                               pinned by
=====================================  =============================================
forces.pairwise / coulomb / images      closed forms (Plummer 2-body, flat self
                                        image, 4-charge sphere+plane), brute force
fields.gun_field / tip_field            closed forms (3x enhancement at the pole)
bem.ellip_ke_complement                 mpmath at 50 digits; DLMF identities
bem.ring_field / ring_potential         brute azimuthal quadrature; finite
                                        differences of the potential; golden vectors
bem.evaluate_axisymmetric_field         golden vectors from the reference
                                        (sphere and post profiles, on/near/far
                                        points); flat disk; grounded sphere
bem.solve                               golden sigma from the reference; grounded
                                        sphere potential
rk45.integrate                          golden runs from the reference (step counts,
                                        events, final state); ballistic closed forms
screens.find_crossings                  cubic with 3 known roots; turnaround; half-
                                        open interval; golden events
tracker.group_indices / setup           sort-by-id KAT; unit round trips
=====================================  =============================================
"""
