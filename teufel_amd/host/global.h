// global.h - constants and exceptions of the host API (the interface of the
// reference's src/global.h:37-44, 78-96, kept alike so code written against
// TEUFEL's classes compiles unchanged against this library).
#pragma once

#include <exception>
#include <string>

#define ElementaryCharge 1.6021766208e-19  // [As] (CODATA-2014, as the reference)
#define mecsquared 0.5109989461e6          // electron rest energy [eV]
#define SpeedOfLight 2.99792458e8          // [m/s]
#define Pi 3.1415926535897932384626433832795029
#define EpsNull 8.854187817e-12            // [As/Vm]
#define MuNull (4 * Pi * 1e-7)             // [Vs/Am]
#define InvRestMass 1.758820025e11         // c^2 / (m c^2) [m^2/s^2/eV]

// Error raised by the library (the reference's IOexception).  The message is
// owned by the exception.
class IOexception : public std::exception {
public:
    explicit IOexception(const std::string& message) : m(message) {}
    const char* what() const throw() override { return m.c_str(); }

private:
    std::string m;
};

// Error for objects used before they are set up (the reference's RANGEexception).
class RANGEexception : public std::exception {
public:
    explicit RANGEexception(const std::string& message) : m(message) {}
    const char* what() const throw() override { return m.c_str(); }

private:
    std::string m;
};
