"""MI355X (gfx950) implementation of the MPI-LES per-time-step hot path.

Modules: layout (reference array layouts), setup_gen (grid, city, masks,
inflow), api (the C ABI of liblesamd.so: reference-signature entry points
and device-resident domains).
"""
