! 1.6-cell S-band gun, r-z boundary (metres): half cell 0-3.25 cm and full
! cell 3.75-10.5 cm of radius 4.2 cm, joined by an iris of radius 1.25 cm
! and thickness 0.5 cm; the beam axis closes the outline.
 &po x=0.0,    y=0.0    &end
 &po x=0.0,    y=0.042  &end
 &po x=0.0325, y=0.042  &end
 &po x=0.0325, y=0.0125 &end
 &po x=0.0375, y=0.0125 &end
 &po x=0.0375, y=0.042  &end
 &po x=0.105,  y=0.042  &end
 &po x=0.105,  y=0.0    &end
 &po x=0.0,    y=0.0    &end
