! DC gun, r-z boundary (metres): a flat cathode at z = 0 (0 V), a 2 cm gap
! to an anode plate at +100 kV with a 2 mm (diameter) bore, a 1 cm drift
! pipe behind it; the outer wall between the electrodes is graded linearly.
 &po x=0.0,   y=0.0,   potential=0 &end
 &po x=0.0,   y=0.010, potential=0 &end
 &po x=0.020, y=0.010, potential=1e5, ramp_potential=1 &end
 &po x=0.020, y=0.001, potential=1e5 &end
 &po x=0.022, y=0.001, potential=1e5 &end
 &po x=0.030, y=0.001, potential=1e5 &end
 &po x=0.030, y=0.0,   potential=1e5 &end
 &po x=0.0,   y=0.0,   potential=0 &end
