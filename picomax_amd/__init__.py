"""PicoMax-MI355X: EPM band structure and chi(q,w) on gfx950."""
