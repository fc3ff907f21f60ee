! TimeStep scheme fragment: include this file in QGN_Module to select RK4.
    integer(c_int), parameter :: time_scheme = QGN_RK4
