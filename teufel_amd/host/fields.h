// fields.h - electromagnetic field value (interface of ElMagField,
// src/fields.h:38-95): E [V/m], B [T], arithmetic and the Poynting vector.
#pragma once

#include "global.h"
#include "vector.h"

class ElMagField {
public:
    ElMagField() {}
    ElMagField(Vector E, Vector B) : vecE(E), vecB(B) {}
    Vector E() const { return vecE; }
    Vector B() const { return vecB; }
    void Zero() { vecE = VectorZero; vecB = VectorZero; }
    // E x B / mu0 [W/m^2]
    Vector Poynting() const { return cross(vecE, vecB) / MuNull; }

    ElMagField operator+(const ElMagField& o) const { return ElMagField(vecE + o.vecE, vecB + o.vecB); }
    ElMagField operator-(const ElMagField& o) const { return ElMagField(vecE - o.vecE, vecB - o.vecB); }
    ElMagField operator*(double f) const { return ElMagField(vecE * f, vecB * f); }
    ElMagField operator/(double f) const { return ElMagField(vecE / f, vecB / f); }
    ElMagField& operator+=(const ElMagField& o) { vecE += o.vecE; vecB += o.vecB; return *this; }
    ElMagField& operator-=(const ElMagField& o) { vecE -= o.vecE; vecB -= o.vecB; return *this; }
    ElMagField& operator*=(double f) { vecE *= f; vecB *= f; return *this; }
    ElMagField& operator/=(double f) { vecE /= f; vecB /= f; return *this; }

private:
    Vector vecE, vecB;
};

#define ElMagFieldZero ElMagField()
