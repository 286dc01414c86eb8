"""MI355X-native force-evaluation hot path of ProtoMol (cell-list cutoff
nonbonded force and PME), hand-written HIP for gfx950 behind a C ABI."""
