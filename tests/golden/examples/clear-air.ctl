# Emitters...
NG = 3
EMITTER[0] = CO2
EMITTER[1] = H2O
EMITTER[2] = O3

# Channels...
ND = 1
NU[0] = 792

# Tables...
TBLBASE = tab/boxcar

# Climatology/Atmosphere
DZ = 1.
ZMAX = 80.
CLIMPATH = ../clim/clim_remedios.tab