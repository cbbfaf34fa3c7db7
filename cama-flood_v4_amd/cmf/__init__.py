"""MI355X-native CaMa-Flood per-substep physics (host side).

Submodules:
  synth   -- seeded synthetic river networks and the reference's input formats
  maps    -- 1-D (NSEQ) map tables built from 2-D maps (host initialisation)
  model   -- device-resident model driving the HIP physics through the C-ABI
  build   -- in-tree build of the native libraries
"""
