"""CPU FP64 oracle of the chatcarlo photon-transport path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package, and only as the checker
(or the timed CPU baseline); the product (chatcarlo_amd) never calls it.

It is a from-scratch NumPy restatement of the reference's algorithms, each
function citing the reference file:line it follows:

* oracle.core      geometry (chatcarlo/geometry.py), cross-section lookup
                   (chatcarlo/materials.py), interaction samplers fed explicit
                   uniforms (chatcarlo/physics.py), h*(10) (dose_coefficients.py),
                   exact segment-voxel overlap (docs/egs5_crosscheck/
                   run_chatcarlo_pdd60.py:49-73) and the stratified track-length
                   estimator (chatcarlo/tally.py:248-293), batch statistics
                   (chatcarlo/tally.py:135-245);
* oracle.source    source and spectrum sampling (chatcarlo/source.py:67-210,
                   chatcarlo/spectrum.py:55-82);
* oracle.transport the round-based batch transport loop
                   (chatcarlo/transport.py:98-312);
* oracle.philox    Philox4x32-10 (Salmon et al., SC'11), the counter-based RNG
                   of the GPU path.

Pinning: see the header of each module; functions with no independent check
say "parity unpinned" there.
"""
