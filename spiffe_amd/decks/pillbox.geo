! closed pillbox cavity, radius 4 cm, length 2.5 cm (metres); the end walls
! lie on the grid edges, where the Neumann left/right boundaries act as walls
 &po x=0.0,   y=0.0  &end
 &po x=0.0,   y=0.04 &end
 &po x=0.025, y=0.04 &end
 &po x=0.025, y=0.0  &end
 &po x=0.0,   y=0.0  &end
