module aeroflux

go 1.21
